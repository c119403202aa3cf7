"""GPU: registered sources and receivers with the time interpolation fused in (SetSourcesInterp,
SetReceiversInterp) equal the reference's flow -- interpadjoint of the coarse wavelet on the
host, injection at the fine rate, extraction of fine traces, interpforward on the host
(src/spacetime.cc:4-127, whose host restatement is pinned to the reference's own build) --
bit for bit, wavefields and coarse traces, for Iso (P) and VTI (P and M)."""
import numpy as np
import pytest

from tests.helpers import host_view, make_model

pytestmark = pytest.mark.gpu

F32 = np.float32


def _setup(kind, shape, fs):
    import wavefd_amd as w
    nx, ny, nz = shape
    cls = {"iso": w.Prop3DAcoIsoDenQ_DEO2_FDTD, "vti": w.Prop3DAcoVTIDenQ_DEO2_FDTD}[kind]
    p = cls(nz=nz, ny=ny, nx=nx, dz=10.0, dy=10.0, dx=10.0, dt=1e-3, nsponge=6, freesurface=fs, freqQ=10.0, qMin=0.5,
            qInterior=80.0)
    for name, arr in make_model(kind, shape, 31).items():
        host_view(p, name)[:] = arr.ravel()
    w.coherence(p, "device")
    w.sync_to_device(p)
    return p


@pytest.mark.parametrize("kind", ["iso", "vti"])
@pytest.mark.parametrize("nh", [1, 3])
@pytest.mark.parametrize("fs", [True, False])
def test_fused_interpolation_equals_host_flow(kind, nh, fs):
    import wavefd_amd as w
    from wavefd_amd import spacetime as st
    shape = (30, 32, 34)  # (nx, ny, nz)
    nx, ny, nz = shape
    dtmod, dtrec = 1e-3, nh * 1e-3
    nd = 23
    nm = (nd - 2) * nh + 1  # the last fine sample the interpolation reaches, plus one past it
    hs = st.interpfilters(dtmod, dtrec, 1)  # wavelet: coarse -> fine
    hr = st.interpfilters(dtmod, dtrec, 0)  # traces: fine -> coarse
    rng = np.random.default_rng(nh)
    wav = rng.standard_normal((1, nd)).astype(F32)
    fsi = 1 if fs else 0
    src = st.hickscoeffs(F32(10.0), F32(10.0), F32(10.0), 0.0, 0.0, 0.0, nz, ny, nx, [52.5], [151.0], [148.5],
                         fs_index=fsi)
    nr = 6
    rcv = st.linearcoeffs(F32(10.0), F32(10.0), F32(10.0), 0.0, 0.0, 0.0, nz, ny, nx, np.full(nr, 33.0),
                          np.linspace(60, 250, nr), np.linspace(70, 230, nr), fs_index=fsi)
    fields = 1 if kind == "iso" else 3
    # the reference's flow: interpolate on the host, run at the fine rate, interpolate back
    a = _setup(kind, shape, fs)
    fine = np.zeros((1, nm), F32)
    st.interpadjoint(hs, fine, wav)
    w.set_sources(a, [src], fine, fields=fields)
    w.set_receivers(a, [rcv], nm)
    for it in range(1, nm + 1):
        w.timestep_inject_extract(a, it)
    traces = w.receiver_data(a, nr, nm)
    want = np.zeros((nr, nd), F32)
    st.interpforward(hr, want, traces)
    # fused
    b = _setup(kind, shape, fs)
    w.set_sources_interp(b, [src], wav, hs, nm, fields=fields)
    w.set_receivers_interp(b, [rcv], nd, hr, nm)
    for it in range(1, nm + 1):
        w.timestep_inject_extract(b, it)
    got = w.receiver_data(b, nr, nd)
    assert np.abs(want).max() > 0
    assert np.array_equal(got, want)
    w.sync_to_host(a)
    w.sync_to_host(b)
    for nm_ in ("pCur", "pOld") + (("mCur", "mOld") if kind != "iso" else ()):
        assert np.array_equal(host_view(a, nm_), host_view(b, nm_)), nm_
    w.free(a)
    w.free(b)


def test_interp_registration_checks():
    import wavefd_amd as w
    from wavefd_amd import spacetime as st
    p = _setup("iso", (20, 20, 20), True)
    h = st.interpfilters(1e-3, 2e-3, 0)
    rcv = st.linearcoeffs(F32(10.0), F32(10.0), F32(10.0), 0.0, 0.0, 0.0, 20, 20, 20, [30.0], [60.0], [70.0],
                          fs_index=1)
    with pytest.raises(RuntimeError, match="nd >= 7"):
        w.set_receivers_interp(p, [rcv], 6, h, 10)
    with pytest.raises(RuntimeError, match="nd >= 7"):
        w.set_sources_interp(p, [rcv], np.zeros((1, 5), F32), h, 10)
    w.free(p)
