"""Input checks of the device ABI, on the CPU: a topology or extents whose
indices would send a kernel outside a device array is refused on the host
(sunhip_validate, which sunhip_grid_create and sunhip_set_extents run
first), with an error naming the bad entry."""
import numpy as np
import pytest

from suntans_amd import cases, device


@pytest.fixture(scope="module")
def st():
    return cases.tiny(1)


def test_valid_grid_passes(st):
    assert device.validate(st) == 0
    assert device.validate(cases.tiny(24, nx=14, ny=9, Nkmax=11, planform="hex", split_every=4, maxfaces=7)) == 0


def _bad(st, name, fn, match):
    a = st.ints(name).copy()
    fn(a)
    with pytest.raises(RuntimeError, match=match):
        device.validate(st, override={name: a})


def test_neigh_out_of_range(st):
    _bad(st, "neigh", lambda a: a.__setitem__(4, st.Nc), "neigh")
    _bad(st, "neigh", lambda a: a.__setitem__(0, -2), "neigh")


def test_face_out_of_range(st):
    _bad(st, "face", lambda a: a.__setitem__(2, st.Ne), "face")
    _bad(st, "face", lambda a: a.__setitem__(2, -1), "face")


def test_grad_out_of_range(st):
    _bad(st, "grad", lambda a: a.__setitem__(1, st.Nc + 3), "grad")
    _bad(st, "grad", lambda a: a.__setitem__(slice(0, 2), -1), "grad")


def test_permutations(st):
    _bad(st, "cellp", lambda a: a.__setitem__(1, a[0]), "cellp")
    _bad(st, "edgep", lambda a: a.__setitem__(3, st.Ne), "edgep")


def test_dists(st):
    _bad(st, "celldist", lambda a: a.__setitem__(2, st.Nc + 1), "celldist")
    _bad(st, "edgedist", lambda a: a.__setitem__(1, a[0] - 1 if a[0] > 0 else -1), "edgedist")


def test_column_sizes(st):
    _bad(st, "Nk", lambda a: a.__setitem__(0, 0), "Nk")
    _bad(st, "Nk", lambda a: a.__setitem__(0, st.scalar("Nkmax") + 1), "Nk")
    _bad(st, "Nkc", lambda a: a.__setitem__(0, -1), "Nkc")


@pytest.mark.parametrize("name,delta,match", [("ctop", 1, "ctop"), ("ctopold", 1, "ctopold"),
                                              ("Nke", 1, "Nke"), ("etop", 1, "etop")])
def test_extents_above_column(st, name, delta, match):
    a = st.ints(name).copy()
    ref = st.ints("Nk" if name.startswith("c") else "Nkc")
    a[3] = ref[3] + delta
    with pytest.raises(RuntimeError, match=match):
        device.validate(st, override={name: a})


@pytest.mark.parametrize("name", ["ctop", "ctopold", "Nke", "etop"])
def test_extents_negative(st, name):
    a = st.ints(name).copy()
    a[2] = -1
    with pytest.raises(RuntimeError, match=name):
        device.validate(st, override={name: a})


def test_grid_create_checks_before_device(st):
    """grid_create refuses a bad topology with the validation message, not
    the missing-device one, so the check needs no GPU."""
    nb = st.ints("neigh").copy()
    nb[0] = st.Nc
    T, keep = device.topology(st, {"neigh": nb})
    import ctypes as C
    h = C.c_void_p()
    rc = device.lib().sunhip_grid_create(C.byref(h), C.byref(T), 0)
    assert rc == -1 and b"neigh" in device.lib().sunhip_error()
