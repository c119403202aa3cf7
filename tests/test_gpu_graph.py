"""GPU: the graph-captured time loop (TimeStepInjectExtractGraph) gives the results of the calls
one by one (TimeStepInjectExtract), bit for bit, when graph segments and eager calls alternate,
with rounds in step on a grid that has them, and it refuses a handle in coherent mode without
touching its state."""
import numpy as np
import pytest

from tests.helpers import host_view, make_fields, make_model, to_oracle

pytestmark = pytest.mark.gpu

F32 = np.float32


def _setup(kind, shape, nt, config=None):
    import wavefd_amd as w
    nx, ny, nz = shape
    cls = {"iso": w.Prop3DAcoIsoDenQ_DEO2_FDTD, "vti": w.Prop3DAcoVTIDenQ_DEO2_FDTD,
           "tti": w.Prop3DAcoTTIDenQ_DEO2_FDTD}[kind]
    prop = cls(nz=nz, ny=ny, nx=nx, dz=10.0, dy=10.0, dx=10.0, dt=1e-3, nsponge=8, freesurface=True, freqQ=10.0,
               qMin=0.5, qInterior=80.0)
    for name, arr in dict(make_model(kind, shape, 3), **make_fields(kind, shape, 3)).items():
        host_view(prop, name)[:] = arr.ravel()
    if config is not None:
        w.set_fused_config(prop, *config)
    w.sync_to_device(prop)
    w.coherence(prop, "device")
    src = w.hickscoeffs(F32(10), F32(10), F32(10), 0.0, 0.0, 0.0, nz, ny, nx, [103.0], [157.5], [163.3])
    wavelet = w.causal_ricker(25.0, 1e-3 * np.arange(nt))[None, :]
    zr = np.full(20, 50.0)
    yr = np.linspace(60, 10 * (ny - 10), 20)
    xr = np.linspace(70, 10 * (nx - 10), 20)
    rcv = w.linearcoeffs(F32(10), F32(10), F32(10), 0.0, 0.0, 0.0, nz, ny, nx, zr, yr, xr)
    w.set_sources(prop, w.source_blocking(src, 4), wavelet, fields=3 if kind != "iso" else 1)
    w.set_receivers(prop, w.receiver_blocking(rcv, 4), nt)
    return prop


def _state(prop, kind, shape, nt):
    import wavefd_amd as w
    w.sync_to_host(prop)
    names = ("pCur", "pOld") + (("mCur", "mOld") if kind != "iso" else ())
    return [to_oracle(prop, n, shape).copy() for n in names] + [w.receiver_data(prop, 20, nt)]


@pytest.mark.parametrize("kind", ["iso", "vti", "tti"])
def test_graph_segments_equal_the_calls_one_by_one(kind):
    import wavefd_amd as w
    shape, nt = (40, 38, 44), 40
    eager = _setup(kind, shape, nt)
    ref13 = None
    for it in range(1, nt + 1):
        w.timestep_inject_extract(eager, it)
        if it == 13:
            ref13 = _state(eager, kind, shape, nt)
    ref = _state(eager, kind, shape, nt)
    g = _setup(kind, shape, nt)
    w.timestep_inject_extract_graph(g, 1, 13)  # an odd count: the buffers end swapped
    for a, b in zip(_state(g, kind, shape, nt), ref13):
        assert np.array_equal(a, b)
    for it in range(14, 21):  # eager calls between the graphs
        w.timestep_inject_extract(g, it)
    w.timestep_inject_extract_graph(g, 21, 7)
    w.timestep_inject_extract_graph(g, 28, 13)
    for a, b in zip(_state(g, kind, shape, nt), ref):
        assert np.array_equal(a, b)
    assert np.abs(ref[-1]).max() > 0


def test_graph_with_rounds_in_step():
    # 45 tiles per plane x 128 chunks of 2 planes: enough workgroups for rounds in step; the
    # round counters are zeroed by a kernel inside the graph
    import wavefd_amd as w
    shape, nt = (256, 128, 128), 3
    eager = _setup("vti", shape, nt, ("64x16", True, 2))
    assert w.fused_config(eager)["lockstep"]
    for it in range(1, nt + 1):
        w.timestep_inject_extract(eager, it)
    ref = _state(eager, "vti", shape, nt)
    g = _setup("vti", shape, nt, ("64x16", True, 2))
    w.timestep_inject_extract_graph(g, 1, nt)
    for a, b in zip(_state(g, "vti", shape, nt), ref):
        assert np.array_equal(a, b)


def test_graph_refuses_coherent_mode():
    import wavefd_amd as w
    shape = (24, 22, 26)
    prop = w.Prop3DAcoIsoDenQ_DEO2_FDTD(nz=shape[2], ny=shape[1], nx=shape[0], nsponge=4)
    w.V(prop)[...] = 1500
    w.B(prop)[...] = 1
    pc = w.PCur(prop)
    pc[...] = np.random.default_rng(0).random(pc.shape, dtype=np.float32)
    before = np.array(pc)
    with pytest.raises(RuntimeError):
        w.timestep_inject_extract_graph(prop, 1, 3)
    assert np.array_equal(np.array(w.PCur(prop)), before)
    w.free(prop)


def test_time_index_out_of_range_is_rejected_before_any_work():
    # checked on the host against the registered lengths: no kernel is launched, no buffer swapped
    import wavefd_amd as w
    shape, nt = (40, 38, 44), 10
    prop = _setup("iso", shape, nt)
    before = _state(prop, "iso", shape, nt)
    for bad in (0, nt + 1):
        with pytest.raises(RuntimeError):
            w.timestep_inject_extract(prop, bad)
    with pytest.raises(RuntimeError):
        w.timestep_inject_extract_graph(prop, nt - 2, 4)
    assert prop._lib.InjectRegistered(prop.p, nt + 1) == -1
    assert prop._lib.ExtractRegistered(prop.p, 0) == -1
    prop._lib.ClearError()
    for a, b in zip(_state(prop, "iso", shape, nt), before):
        assert np.array_equal(a, b)
