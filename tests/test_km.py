"""Kubelka-Munk mix on MFMA against the FP64 restatement: |dR| <= 1e-4 on
BF16 concentrations and FP32 K, S tables; and the reference form (16 bands,
tint-weighted pigment table of src/color/mod.rs:243-258) against layer_ks
and km_layer at large thickness (shaders/km.wgsl:1-33)."""
import ctypes as C

import numpy as np
import pytest

import canvaz_film as czf
import km
import oracle as orc


def bf16_bits(x):
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def km_inputs(cells, bands, seed):
    rng = np.random.default_rng(seed)
    conc = bf16_bits(rng.dirichlet(np.full(12, 0.3), cells))
    kt = np.exp(rng.uniform(np.log(0.01), np.log(10), (12, bands))).astype(np.float32)
    st = np.exp(rng.uniform(np.log(0.01), np.log(10), (12, bands))).astype(np.float32)
    return conc, kt, st


def oracle_mix(conc, kt, st):
    cells, bands = conc.shape[0], kt.shape[1]
    out = np.zeros((cells, bands), np.float64)
    orc.lib().of_km_mix_f64(conc.ctypes.data_as(C.POINTER(C.c_uint16)), kt.ctypes.data_as(C.POINTER(C.c_float)),
                            st.ctypes.data_as(C.POINTER(C.c_float)), cells, bands,
                            out.ctypes.data_as(C.POINTER(C.c_double)))
    return out


def test_prepare_splits_tables():
    _, kt, st = km_inputs(1, 36, 1)
    ops = km.prepare(kt, st).reshape(6, 64, 8)
    f = (ops.astype(np.uint32) << 16).view(np.float32)
    for t in range(3):
        for lane in range(64):
            col, grp = lane & 15, lane >> 4
            band = 16 * t + col
            for j in range(8):
                k = 8 * grp + j
                if band >= 36 or 12 <= k < 16 or k >= 28:
                    assert ops[t, lane, j] == 0
    # hi + lo reproduces the table to 2^-16 relative
    rec = np.zeros((12, 36), np.float64)
    for lane in range(64):
        col, grp = lane & 15, lane >> 4
        for t in range(3):
            band = 16 * t + col
            if band >= 36:
                continue
            for j in range(8):
                k = 8 * grp + j
                if k < 12:
                    rec[k, band] += f[t, lane, j]
                elif 16 <= k < 28:
                    rec[k - 16, band] += f[t, lane, j]
    assert np.allclose(rec, kt, rtol=2 ** -15)


@pytest.mark.gpu
@pytest.mark.parametrize("bands,cells", [(36, 100003), (16, 4099), (48, 777)])
def test_km_mix_vs_fp64(gpu, bands, cells):
    import torch
    conc, kt, st = km_inputs(cells, bands, seed=bands)
    ops = torch.from_numpy(km.prepare(kt, st).view(np.int16)).cuda()
    c = torch.from_numpy(conc.view(np.int16)).cuda().view(torch.bfloat16).reshape(cells, 12)
    out = km.mix(c, ops, bands)
    torch.cuda.synchronize()
    ref = oracle_mix(conc, kt, st)
    err = np.abs(out.cpu().numpy().astype(np.float64) - ref)
    assert err.max() <= 1e-4, err.max()
    assert np.isfinite(out.cpu().numpy()).all()


@pytest.mark.gpu
def test_km_mix_reference_form(gpu):
    """16 bands of the default pigment table, tint folded into the table: the
    ratio K/S of layer_ks does not depend on its amount/weight factor."""
    import torch
    pt = czf.default_pigment_table().reshape(12, 40)
    tint = pt[:, 34:35]
    kt = (pt[:, :16] * tint).astype(np.float32)
    stt = (pt[:, 16:32] * tint).astype(np.float32)
    rng = np.random.default_rng(3)
    cells = 2048
    conc = bf16_bits(rng.dirichlet(np.full(12, 0.3), cells))
    ops = torch.from_numpy(km.prepare(kt, stt).view(np.int16)).cuda()
    c = torch.from_numpy(conc.view(np.int16)).cuda().view(torch.bfloat16).reshape(cells, 12)
    out = km.mix(c, ops, 16).cpu().numpy()
    o = orc.OracleFilm(4, 4, czf.default_settings(), czf.default_pigment_table(), czf.cmf_table())
    cf = (conc.astype(np.uint32) << 16).view(np.float32)
    for cell in range(0, cells, 97):
        pig = np.zeros(16, np.float32)
        pig[:12] = cf[cell]
        for b in range(16):
            k, s = C.c_float(), C.c_float()
            orc.lib().of_layer_ks(o._p, pig.ctypes.data_as(C.POINTER(C.c_float)), b, C.byref(k), C.byref(s))
            R, T = C.c_float(), C.c_float()
            orc.lib().of_km_layer(k.value, s.value, 1.0, C.byref(R), C.byref(T))
            assert abs(out[cell, b] - R.value) <= 1e-4, (cell, b, out[cell, b], R.value)
