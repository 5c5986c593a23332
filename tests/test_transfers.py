"""Batched host<->device transfers (cable_gpu_upload_fields and
cable_gpu_download_fields): the same bits as the per-field calls, for every
kind of field, with a tile order set and at both precisions."""
import numpy as np
import pytest

from cable_amd.fields import BY_NAME, FIELDS

pytestmark = pytest.mark.gpu

NAMES = ["veg%iveg", "ssnow%isflag", "ssnow%tgg", "ssnow%wb", "met%fsd", "canopy%fe", "canopy%fwsoil",
         "rad%albedo", "soil%zse_vec", "bgc%cplant"]


def _values(rng, name, mp):
    f = BY_NAME[name]
    if f.prec == "I":
        return rng.integers(1, 17, size=f.shape(mp)).astype(np.float64)
    return rng.normal(size=f.shape(mp))


@pytest.mark.parametrize("precision", [0, 1])
def test_batched_transfers_match_per_field_transfers(precision):
    from cable_amd import gpu
    print(f"[transfers] precision {precision}", flush=True)
    mp = 5003
    rng = np.random.default_rng(precision)
    assert all(n in BY_NAME for n in NAMES)
    g = gpu.CableGPU(mp, precision)
    g.set_tile_order(rng.permutation(mp))
    a = {n: _values(rng, n, mp) for n in NAMES}
    for n in NAMES:
        g.upload(n, a[n])
    got = g.download_fields(NAMES)
    for n in NAMES:
        np.testing.assert_array_equal(got[n], g.download(n), err_msg=n)
    # FP64 contexts keep every value; mixed ones round default-REAL fields
    # to FP32 on the device, both ways alike
    if precision == 0:
        for n in NAMES:
            np.testing.assert_array_equal(got[n], a[n], err_msg=n)
    g.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_batched_upload_equals_per_field_upload(precision):
    from cable_amd import gpu
    print(f"[transfers] upload, precision {precision}", flush=True)
    mp = 4099
    rng = np.random.default_rng(10 + precision)
    order = rng.permutation(mp)
    vals = {n: _values(rng, n, mp) for n in NAMES}
    out = []
    for batched in (True, False):
        g = gpu.CableGPU(mp, precision)
        g.set_tile_order(order)
        if batched:
            g.upload_fields(vals)
        else:
            for n in NAMES:
                g.upload(n, vals[n])
        out.append({n: g.download(n) for n in NAMES})
        g.close()
    for n in NAMES:
        np.testing.assert_array_equal(out[0][n], out[1][n], err_msg=n)


def test_every_field_round_trips_in_one_batch():
    """All fields at once (the drop-in's first call) through one transfer
    each way."""
    from cable_amd import gpu
    print("[transfers] all fields", flush=True)
    mp = 1024
    rng = np.random.default_rng(3)
    names = [f.name for f in FIELDS]
    vals = {n: _values(rng, n, mp) for n in names}
    g = gpu.CableGPU(mp, 0)
    g.upload_fields(vals)
    back = g.download_fields(names)
    g.close()
    for n in names:
        np.testing.assert_array_equal(back[n], vals[n], err_msg=n)


def test_batched_transfers_refuse_bad_arguments():
    """A field id out of range or an unknown host type is refused with a
    message, before anything moves; an empty batch is a no-op."""
    import ctypes
    from cable_amd import gpu
    print("[transfers] bad arguments", flush=True)
    g = gpu.CableGPU(64, 0)
    L = gpu.lib()
    a = np.zeros(64)
    ptrs = (ctypes.c_void_p * 1)(a.ctypes.data)
    for fid, ht in ((10000, gpu.HOST_F64), (-1, gpu.HOST_F64), (BY_NAME["canopy%fe"].fid, 7)):
        ids = np.array([fid], dtype=np.int32)
        types = np.array([ht], dtype=np.int32)
        for fn in (L.cable_gpu_upload_fields, L.cable_gpu_download_fields):
            assert fn(g.h, 1, ids.ctypes.data, ptrs, types.ctypes.data) != 0
            assert b"fields" in L.cable_gpu_last_error()
    g.upload_fields({})
    assert g.download_fields([]) == {}
    g.close()

