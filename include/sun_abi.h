/*
 * sun_abi.h -- binary layout of the SUNTANS state structs that the drop-in
 * solver library reads.
 *
 * The reference passes its grid, physical state and run properties to the
 * elliptic solvers and to scalar transport as three structs (gridT, physT,
 * propT) whose fields are read by offset.  A library that replaces those
 * functions at link time therefore has to see the very same layout: every
 * field below has the type and position of its counterpart in
 *   main/grid.h:32-149  (gridT)
 *   main/phys.h:35-177  (physT)
 *   main/phys.h:183-203 (propT)
 * and the serial MPI stand-ins of main/no-mpi.h:25-29 (numprocs=1).
 * tests/test_abi_layout.py compiles both headers and compares sizeof and
 * every offsetof.
 *
 * When a translation unit has already included the reference's own grid.h
 * and phys.h (their include guards are _grid_h and _phys_h), the definitions
 * here are skipped and the reference's are used unchanged.
 */
#ifndef SUN_ABI_H
#define SUN_ABI_H

#include <stdio.h>

#ifndef REAL
#define REAL double
#endif

#if !defined(_phys_h) && !defined(_grid_h)

#if !defined(_no_mpi_h) && !defined(MPI_VERSION)
/* Serial build: communicator, status and request are plain ints. */
typedef int MPI_Comm;
typedef int MPI_Status;
typedef int MPI_Request;
#endif

typedef struct _gridT {
  /* planform geometry: points, Voronoi points, edge centres, depth */
  REAL *xp, *yp, *xv, *yv, *xe, *ye, *dv;
  /* vertical spacing: reference levels, cell and flux-face heights */
  REAL *dz;
  REAL **dzz;
  REAL **dzf;
  REAL *dzfB;
  REAL **dzzold;
  REAL *dztop, *dzbot;
  REAL dzsmall;
  /* connectivity */
  int *face, *edges, *cells, *neigh, *eneigh, *grad, *gradf, *mark, *edge_id, *normal;
  /* nodal neighbour tables */
  int *localtoglobalpoints;
  int *numppneighs;  int **ppneighs;
  int *numpeneighs;  int **peneighs;
  int *numpcneighs;  int **pcneighs;
  REAL **Actotal;
  /* partition graph */
  int *xadj, *adjncy, *vtxdist, *cellbnddist, *edgebnddist;
  /* computational ordering */
  int *celldist, *edgedist, *cellp, *edgep;
  /* halo tables */
  int **cell_send, **cell_recv, **edge_send, **edge_recv;
  int *num_cells_send, *num_cells_recv, *num_cells_send_first, *num_cells_recv_first;
  int *num_edges_send, *num_edges_recv;
  int maxtosend, maxtorecv;
  MPI_Status *status;
  MPI_Request *request;
  REAL **recv, **send;
  int *total_cells_send, *total_cells_recv, *total_cells_sendW, *total_cells_recvW;
  int *total_edges_send, *total_edges_recv;
  /* global/local maps */
  int *mnptr, *eptr, *part, *order, *epart, *vwgt;
  /* metrics */
  REAL *df, *dg, *def, *Ac, *n1, *n2, *xi;
  int *numneighs;
  int **neighs;
  int Nneighs;
  int *myneighs;
  /* vertical extents */
  int *Nk, *ctop, *ctopold, *etop, *etopold, *Nkp, *Nke, *Nkc;
  int Nkmax, Nc, Nc_comp, Ne, Ne_comp, Np, Nge;
  int *nfaces;
  int maxfaces, stairstep, fixdzz;
  REAL smoothbot;
} gridT;

typedef enum _interpolation { nRT1, nRT2, tRT1, tRT2, QUAD, PEROT, LSQ } interpolation;

typedef struct _physT {
  REAL **u, **uc, **vc, **wc, **wf;
  REAL ***nRT1u, ***nRT1v;
  REAL **nRT2u, **nRT2v, **tRT1, **tRT2;
  REAL **uold, **vold;
  REAL *D;
  REAL **w, **q, **qc, **s, **T, **s0, **rho;
  REAL *h, *hcorr;
  unsigned char *active;
  REAL **boundary_u, **boundary_v, **boundary_w, **boundary_s, **boundary_T,
       **boundary_tmp, **boundary_rho;
  REAL *boundary_h, *boundary_flag;
  REAL **nu_tv, **kappa_tv, **nu_lax;
  REAL *tau_T, *tau_B, *CdT, *CdB;
  REAL **qT, **lT;
  REAL mass, mass0, volume, volume0, Ep, Ep0, Ek, Eflux1, Eflux2, Eflux3, Eflux4, smin, smax;
  /* free-surface solver work arrays */
  REAL *htmp, *hold, *htmp2, *htmp3, *hcoef, *hfcoef;
  /* 3-D work arrays */
  REAL **stmp, **stmp2, **stmp3, **utmp, **utmp2, **ut;
  REAL **Cn_R, **Cn_T, **Cn_U, **Cn_U2, **Cn_W, **Cn_W2, **Cn_q, **Cn_l;
  REAL **wnew, **wtmp, **wtmp2, **qtmp;
  REAL **TP, **TB, **TD;
  /* per-column scratch, Nkmax+1 long */
  REAL *ap, *am, *bp, *bm;
  REAL *a, *b, *c, *d;
  /* horizontal face scalars */
  REAL **SfHp, **SfHm;
  /* vertical TVD scratch */
  REAL *Cp, *Cm, *rp, *rm, *wp, *wm;
  REAL **gradSx, **gradSy;
  /* heat flux */
  REAL *Tsurf, *dT;
  REAL **Ttmp;
  /* output staging */
  REAL *tmpvar, *tmpvarW, *tmpvarE;
  /* least-squares interpolation */
  REAL **A, **Apr, **AT;
  REAL *bpr;
} physT;

typedef struct _propT {
  REAL dt, Cmax, rtime, amp, omega, flux, timescale, theta0, theta, thetaM,
       thetaS, thetaB, nu, nu_H, tau_T, z0T, CdT, z0B, CdB, CdW, relax,
       epsilon, qepsilon, resnorm, dzsmall, beta, kappa_s, kappa_sH, gamma,
       kappa_T, kappa_TH, grav, Coriolis_f, CmaxU, CmaxW, laxWendroff_Vertical,
       latitude;
  int ntout, ntoutStore, ntprog, nsteps, nstart, n, ntconserve, nonhydrostatic,
      cgsolver, maxiters, qmaxiters, hprecond, qprecond, volcheck, masscheck,
      nonlinear, linearFS, newcells, wetdry, sponge_distance, sponge_decay,
      thetaramptime, readSalinity, readTemperature, turbmodel, TVD, horiTVD,
      vertTVD, TVDsalt, TVDtemp, TVDturb, laxWendroff, stairstep, AB,
      TVDmomentum, conserveMomentum, mergeArrays, computeSediments;
  FILE *FreeSurfaceFID, *HorizontalVelocityFID, *VerticalVelocityFID,
       *SalinityFID, *BGSalinityFID, *InitSalinityFID, *InitTemperatureFID,
       *TemperatureFID, *PressureFID, *VerticalGridFID, *ConserveFID,
       *StoreFID, *StartFID, *EddyViscosityFID, *ScalarDiffusivityFID;
  interpolation interp;
  int prettyplot;
  int metmodel, varmodel, outputNetcdf, metncid, netcdfBdy, netcdfBdyFileID,
      readinitialnc, initialNCfileID, calcage, agemethod, calcaverage;
  int outputNetcdfFileID, averageNetcdfFileID;
  REAL nctime, toffSet, gmtoffset;
  int nctimectr, avgtimectr, avgctr, avgfilectr, ntaverage, nstepsperncfile, ncfilectr;
  REAL nugget, sill, range, Lsw, Cda, Ce, Ch;
  char starttime[15], basetime[15];
} propT;

#endif /* reference headers not included */

#endif /* SUN_ABI_H */
