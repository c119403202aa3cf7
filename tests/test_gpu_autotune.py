"""GPU: the fused step's launch configuration (tile, rounds in step, x-chunk length) changes no
result, and the autotuner leaves the handle's state as it found it and keeps its choice."""
import numpy as np
import pytest

from tests.helpers import make_fields, make_model

pytestmark = pytest.mark.gpu


def _prop(kind, shape, seed=3, fs=True):
    import wavefd_amd as w
    nx, ny, nz = shape
    cls = {"iso": w.Prop3DAcoIsoDenQ_DEO2_FDTD, "vti": w.Prop3DAcoVTIDenQ_DEO2_FDTD,
           "tti": w.Prop3DAcoTTIDenQ_DEO2_FDTD}[kind]
    p = cls(nz=nz, ny=ny, nx=nx, dz=10.0, dy=10.0, dx=10.0, dt=1e-3, nsponge=8, freesurface=fs)
    getters = {"v": w.V, "b": w.B, "eps": w.Eps, "eta": w.Eta, "f": w.F, "sinTheta": w.SinTheta,
               "cosTheta": w.CosTheta, "sinPhi": w.SinPhi, "cosPhi": w.CosPhi, "pCur": w.PCur, "pOld": w.POld,
               "mCur": w.MCur, "mOld": w.MOld}
    arrs = dict(make_model(kind, shape, seed), **make_fields(kind, shape, seed))
    for k, a in arrs.items():
        getters[k](p)[...] = a.T  # oracle (nx, ny, nz) C order == Julia (nz, ny, nx) column major
    return p


def _run(kind, shape, config, nsteps=4):
    import wavefd_amd as w
    from wavefd_amd import tuning as at
    p = _prop(kind, shape)
    if config is not None:
        at.set_fused_config(p, *config)
    eff = at.fused_config(p)
    for _ in range(nsteps):
        w.propagateforward(p)
    out = [np.array(w.PCur(p)), np.array(w.POld(p)), np.array(w.PSpace(p))]
    if kind != "iso":
        out += [np.array(w.MCur(p)), np.array(w.MOld(p))]
    w.free(p)
    return eff, out


@pytest.mark.parametrize("kind", ["iso", "vti", "tti"])
def test_results_do_not_depend_on_the_configuration(kind):
    shape = (80, 40, 70)
    _, ref = _run(kind, shape, None)
    for config in (("64x16", False, 0), ("32x32", False, 0), ("64x16", False, 27), ("32x32", False, 14),
                   ("64x16", False, 80), ("wide", False, 0), ("wide", False, 9), ("wide2", False, 0),
                   ("wide2", False, 11)):
        eff, got = _run(kind, shape, config)
        # "wide2" is the plain Iso step's; the other physics take "wide" for it
        assert eff["tile"] == (config[0] if kind == "iso" or config[0] != "wide2" else "wide")
        if config[2]:
            assert eff["chunk"] == config[2]
        for a, b in zip(ref, got):
            assert np.array_equal(a, b), (config, eff)


@pytest.mark.parametrize("kind", ["vti", "tti"])
@pytest.mark.parametrize("tile,chunk", [("64x16", 2), ("wide", 1)])
def test_rounds_in_step_change_no_result(kind, tile, chunk):
    # enough workgroups for 20 rounds of one per CU (5120 on 256 CUs; 12 rounds on the wide tile): 45 tiles per plane
    # (128 x 128, 64 x 16 tile) times 128 chunks of 2 x-planes
    shape = (256, 128, 128)
    # (wide tile: 24 tiles per plane times 256 chunks of one x-plane)
    eff_off, ref = _run(kind, shape, (tile, False, chunk), nsteps=2)
    eff_on, got = _run(kind, shape, (tile, True, chunk), nsteps=2)
    assert not eff_off["lockstep"]
    assert eff_on["lockstep"], eff_on
    for a, b in zip(ref, got):
        assert np.array_equal(a, b)


def test_set_fused_config_rejects_bad_values():
    import wavefd_amd as w
    p = _prop("iso", (24, 20, 22))
    for bad in ((5, 0, 0), (-1, 0, 0), (0, 2, 0), (0, 0, -1)):
        assert p._lib.SetFusedConfig(p.p, *bad) == -1
        p._lib.ClearError()
    w.free(p)


@pytest.mark.parametrize("kind", ["iso", "tti"])
@pytest.mark.parametrize("mode", ["coherent", "device"])
def test_autotune_restores_state_and_keeps_the_best(kind, mode):
    import wavefd_amd as w
    from wavefd_amd import tuning as at
    shape = (60, 36, 50)
    p = _prop(kind, shape)
    w.coherence(p, mode)
    if mode == "device":
        w.sync_to_device(p)
    views = [w.PCur(p), w.POld(p)] + ([w.MCur(p), w.MOld(p)] if kind != "iso" else [])
    before = [np.array(v) for v in views]
    best, table = at.autotune(p, nsteps=3, warmup=1)
    assert len(table) >= 2
    assert best["ms"] == min(r[3] for r in table)
    eff = at.fused_config(p)
    assert (eff["tile"], eff["lockstep"], eff["chunk"]) == (best["tile"], best["lockstep"], best["chunk"])
    if mode == "device":
        w.sync_to_host(p)
    for v, b in zip(views, before):
        assert np.array_equal(np.array(v), b)
    # the tuned handle continues exactly like an untuned one
    w.propagateforward(p)
    if mode == "device":
        w.sync_to_host(p)
    _, ref = _run(kind, shape, None, nsteps=1)
    assert np.array_equal(np.array(w.PCur(p)), ref[0])
    w.free(p)


def test_tilesize_mirrors_the_reference_signature():
    import wavefd_amd as w
    rngx, rngy, speeds = w.tilesize(w.Prop3DAcoTTIDenQ_DEO2_FDTD, nz=40, ny=30, nx=36, rngy=[4], rngx=[4, 8], N=2)
    assert list(rngx) == [4, 8] and list(rngy) == [4]
    assert speeds.shape == (1, 2)
    assert np.all(np.isfinite(speeds)) and np.all(speeds > 0)
