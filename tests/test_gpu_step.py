"""GPU parity of the time step (fused and two-pass) against the CPU oracle.

The HIP kernels evaluate the reference's expressions in the reference's order without FMA
contraction, so single steps must match the oracle bit for bit; multi-step runs are held to
the same bar and, as a backstop, to relative L2 <= 1e-6.
"""
import numpy as np
import pytest

from tests.helpers import MODEL, host_view, make_fields, make_model, rel_l2, to_oracle, ulp_diff

pytestmark = pytest.mark.gpu

# the last two: the last tile along z and y writes the grid's last (annulus) cells from its upper
# halo (ny, nz a few cells past a multiple of the outputs per tile: 57 x 17 for the VTI/TTI wide
# tile, 57 x 9 for 64 x 16 and wide2, 25 x 25 for 32 x 32, 57 x 25 for Iso's wide)
SHAPES = [(37, 41, 45), (20, 23, 61), (19, 60, 130), (21, 37, 60), (23, 28, 53)]


def _pair(oracle, kind, shape, fs, seed, nsponge=6):
    import wavefd_amd as w
    nx, ny, nz = shape
    d = (10.0, 12.5, 8.0)
    cls = {"iso": w.Prop3DAcoIsoDenQ_DEO2_FDTD, "vti": w.Prop3DAcoVTIDenQ_DEO2_FDTD,
           "tti": w.Prop3DAcoTTIDenQ_DEO2_FDTD}[kind]
    ocls = {"iso": oracle.OracleIso, "vti": oracle.OracleVTI, "tti": oracle.OracleTTI}[kind]
    prop = cls(nz=nz, ny=ny, nx=nx, dz=d[2], dy=d[1], dx=d[0], dt=1e-3, nsponge=nsponge, freesurface=fs,
               freqQ=10.0, qMin=0.5, qInterior=80.0)
    orc = ocls(nx, ny, nz, *d, dt=1e-3, fs=fs, nsponge=nsponge, freqQ=10.0, qMin=0.5, qInterior=80.0, nthread=8)
    model = make_model(kind, shape, seed)
    fields = make_fields(kind, shape, seed)
    for name, arr in list(model.items()) + list(fields.items()):
        host_view(prop, name)[:] = arr.ravel()
        getattr(orc, name)[...] = arr
    return prop, orc


def _compare(prop, orc, kind, shape, names, exact=True):
    worst = {}
    for name in names:
        got = to_oracle(prop, name, shape)
        ref = getattr(orc, name)
        worst[name] = (ulp_diff(got, ref), rel_l2(got, ref))
        if exact:
            assert np.array_equal(got, ref), (name, worst[name])
        else:
            assert worst[name][1] <= 1e-6, (name, worst[name])
    return worst


@pytest.mark.parametrize("kind", ["iso", "vti", "tti"])
@pytest.mark.parametrize("fs", [False, True])
def test_q_profile_matches_oracle(oracle, kind, fs):
    prop, orc = _pair(oracle, kind, (21, 19, 23), fs, 1, nsponge=5)
    got = to_oracle(prop, "q", (21, 19, 23))
    assert np.array_equal(got, orc.q)


@pytest.mark.parametrize("kind", ["iso", "vti", "tti"])
@pytest.mark.parametrize("fs", [False, True])
def test_unphysical_q(oracle, kind, fs):
    """qMin = qInterior = 0 (src/absorb.h:105-109, test/absorb.jl:85-99): a new handle's Q is all
    zero; set up again over a profile, only the first two values are cleared, as the reference's
    memset of sizeof(float *) bytes does -- the oracle's restatement of it, pinned against the
    reference's own absorb.h.  The step then reads that array."""
    import wavefd_amd as w
    shape = (21, 19, 23)
    nx, ny, nz = shape
    cls = {"iso": w.Prop3DAcoIsoDenQ_DEO2_FDTD, "vti": w.Prop3DAcoVTIDenQ_DEO2_FDTD,
           "tti": w.Prop3DAcoTTIDenQ_DEO2_FDTD}[kind]
    prop = cls(nz=nz, ny=ny, nx=nx, dz=8.0, dy=12.5, dx=10.0, dt=1e-3, nsponge=5, freesurface=fs, freqQ=10.0,
               qMin=0.0, qInterior=0.0)
    assert not to_oracle(prop, "q", shape).any()
    prop._lib.SetupDtOmegaInvQ(prop.p, 10.0, 0.5, 80.0)
    want = oracle.setup_q(fs, nx, ny, nz, 5, 1e-3, 10.0, 0.5, 80.0)
    assert np.array_equal(to_oracle(prop, "q", shape), want)
    prop._lib.SetupDtOmegaInvQ(prop.p, 10.0, 0.0, 0.0)
    oracle.setup_q(fs, nx, ny, nz, 5, 1e-3, 10.0, 0.0, 0.0, out=want)
    got = to_oracle(prop, "q", shape)
    assert np.array_equal(got, want) and not got.ravel()[:2].any() and got.ravel()[2:].all()
    w.free(prop)


@pytest.mark.parametrize("kind", ["iso", "vti", "tti"])
@pytest.mark.parametrize("fs", [False, True])
@pytest.mark.parametrize("fused", ["64x16", "32x32", "wide", "wide2", None])
@pytest.mark.parametrize("shape", SHAPES)
def test_steps_bit_exact(oracle, kind, fs, fused, shape):
    import wavefd_amd as w
    prop, orc = _pair(oracle, kind, shape, fs, 7)
    w.set_fused_step(prop, fused is not None)
    if fused:
        w.set_fused_tile(prop, fused)
    for _ in range(5):
        w.propagateforward(prop)
        orc.time_step()
    names = ["pCur", "pOld", "pSpace"] + (["mCur", "mOld", "mSpace"] if kind != "iso" else [])
    _compare(prop, orc, kind, shape, names)


@pytest.mark.parametrize("kind", ["iso", "vti", "tti"])
@pytest.mark.parametrize("chunk", [5, 13])
@pytest.mark.parametrize("tile", ["wide", "wide2"])
def test_wide_tile_short_chunks_bit_exact(oracle, kind, chunk, tile):
    # several x-chunks per grid, each padded to whole 8-step rounds of the wide kernel's loop
    import wavefd_amd as w
    shape = (37, 41, 70)
    prop, orc = _pair(oracle, kind, shape, True, 5)
    w.set_fused_config(prop, tile, False, chunk)
    for _ in range(3):
        w.propagateforward(prop)
        orc.time_step()
    names = ["pCur", "pOld", "pSpace"] + (["mCur", "mOld", "mSpace"] if kind != "iso" else [])
    _compare(prop, orc, kind, shape, names)


@pytest.mark.parametrize("fs", [False, True])
@pytest.mark.parametrize("tile", ["auto", "wide", "64x16", "32x32"])
@pytest.mark.parametrize("shape", [(30, 29, 33), (21, 37, 60)])
def test_vti_linear_step_bit_exact(oracle, fs, tile, shape):
    import wavefd_amd as w
    prop, orc = _pair(oracle, "vti", shape, fs, 3)
    w.set_fused_tile(prop, tile)
    for _ in range(4):
        w.propagateforward_linear(prop)
        orc.time_step_linear()
    _compare(prop, orc, "vti", shape, ["pCur", "pOld", "mCur", "mOld"])


@pytest.mark.parametrize("kind", ["iso", "vti", "tti"])
def test_many_steps_device_mode(oracle, kind):
    """200 steps in the stream-ordered device mode, then compare."""
    import wavefd_amd as w
    shape = (40, 36, 50)
    prop, orc = _pair(oracle, kind, shape, True, 11)
    w.coherence(prop, "device")
    w.sync_to_device(prop)
    for _ in range(200):
        w.propagateforward(prop)
        orc.time_step()
    w.sync_to_host(prop)
    names = ["pCur", "pOld"] + (["mCur", "mOld"] if kind != "iso" else [])
    _compare(prop, orc, kind, shape, names, exact=False)


def _prop_only(kind, shape, seed):
    import wavefd_amd as w
    nx, ny, nz = shape
    cls = {"iso": w.Prop3DAcoIsoDenQ_DEO2_FDTD, "vti": w.Prop3DAcoVTIDenQ_DEO2_FDTD,
           "tti": w.Prop3DAcoTTIDenQ_DEO2_FDTD}[kind]
    prop = cls(nz=nz, ny=ny, nx=nx, dz=8.0, dy=12.5, dx=10.0, dt=1e-3, nsponge=6, freesurface=True, freqQ=10.0,
               qMin=0.5, qInterior=80.0)
    for name, arr in list(make_model(kind, shape, seed).items()) + list(make_fields(kind, shape, seed).items()):
        host_view(prop, name)[:] = arr.ravel()
    return prop


@pytest.mark.parametrize("kind", ["iso", "vti", "tti"])
def test_q_table_and_q_array_give_the_same_step(kind):
    # while dtOmegaInvQ is the SetupDtOmegaInvQ profile the fused step reads Q from a table by
    # sponge distance; once Q's device address is handed out it reads the array: same bits
    import wavefd_amd as w
    shape = (40, 33, 37)
    a, b = _prop_only(kind, shape, 11), _prop_only(kind, shape, 11)
    w.coherence(a, "device")
    w.coherence(b, "device")
    w.sync_to_device(a)
    w.sync_to_device(b)
    assert w.device_ptr(b, "dtOmegaInvQ")  # b: the array path from here on
    for _ in range(4):
        w.propagateforward(a)
        w.propagateforward(b)
    w.sync_to_host(a)
    w.sync_to_host(b)
    names = ["pCur", "pOld", "pSpace"] + (["mCur", "mOld", "mSpace"] if kind != "iso" else [])
    for name in names:
        assert np.array_equal(to_oracle(a, name, shape), to_oracle(b, name, shape)), name


@pytest.mark.parametrize("kind", ["iso", "tti"])
def test_written_q_is_used(oracle, kind):
    # Q written through its host view (a caller's own attenuation model) is what the step uses
    import wavefd_amd as w
    shape = (37, 41, 45)
    prop, orc = _pair(oracle, kind, shape, True, 13)
    q = (0.002 + 0.01 * np.random.default_rng(5).random(shape)).astype(np.float32)
    view = host_view(prop, "q")
    # a later SetupDtOmegaInvQ does not bring the table back: the view handed out may still be
    # written, and is
    prop._lib.SetupDtOmegaInvQ(prop.p, 10.0, 0.5, 80.0)
    view[:] = q.ravel()
    orc.q[...] = q
    for _ in range(3):
        w.propagateforward(prop)
        orc.time_step()
    names = ["pCur", "pOld", "pSpace"] + (["mCur", "mOld", "mSpace"] if kind != "iso" else [])
    _compare(prop, orc, kind, shape, names)
