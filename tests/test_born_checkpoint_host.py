"""Host logic of the Born driver's checkpointed and predictive snapshot stores (no GPU): the segment slots of
every step and wavefield are distinct and inside the segment buffer, checkpoints are taken
before steps 1, k+1, 2k+1, ..., the coupled forward pass writes the last segment straight into
the segment buffer, a separate Born pass runs the background segment by segment with no
recomputation, and the adjoint recomputes every segment but the last one once."""
import types

import numpy as np
import pytest

from wavefd_amd import born


class _Buf:
    next_ptr = 1 << 20

    def __init__(self, nbytes):
        self.nbytes, self.ptr = nbytes, _Buf.next_ptr
        _Buf.next_ptr += nbytes + (1 << 20)

    def free(self):
        pass


@pytest.fixture
def store_factory(monkeypatch):
    copies = []
    monkeypatch.setattr(born.ginsu, "DeviceBuffer", _Buf)
    monkeypatch.setattr(born.ginsu, "copy", lambda dst, src, n, stream=None: copies.append((dst, src, n)))
    monkeypatch.setattr(born.prop3d, "device_ptr", lambda prop, name: hash(name) & 0xFFFF)

    def make(nt, nf, k, ncell=1000):
        run = types.SimpleNamespace(nf=nf, nt=nt, ncell=ncell, prop=None, stream=None)
        st = born._CheckpointStore(run, k, True)
        recomputed = []
        st._recompute = lambda j, first=False: (recomputed.append((j, first)), setattr(st, "cached", j))
        return st, copies, recomputed

    return make


@pytest.mark.parametrize("nt,k", [(45, 1), (45, 7), (45, 16), (45, 45), (45, 50), (200, 20)])
@pytest.mark.parametrize("nf", [1, 2])
@pytest.mark.parametrize("segmented", [False, True])
def test_slots_checkpoints_and_recompute_schedule(store_factory, nt, k, nf, segmented):
    st, copies, recomputed = store_factory(nt, nf, k)
    nseg = (nt + k - 1) // k
    slots = {st._slot(kind, n) for n in range(1, nt + 1) for kind in range(nf)
             if (n - 1) // k == (nt - 1) // k}
    seg = [st._slot(kind, n) for n in range(1, min(k, nt) + 1) for kind in range(nf)]
    assert len(set(seg)) == len(seg)
    assert min(seg) == st.seg.ptr and max(seg) + 4 * st.ncell == st.seg.ptr + st.seg.nbytes
    if not segmented:
        # coupled forward: checkpoints before steps 1, k+1, ...; the last segment goes to the
        # segment buffer
        for n in range(1, nt + 1):
            st.begin(n)
            for kind in range(nf):
                o = st.out(kind, n)
                assert (o in slots) == ((n - 1) // k == nseg - 1)
        assert len(copies) == 2 * nf * nseg
        ck = sorted({c[0] for c in copies})
        assert len(ck) == 2 * nf * nseg and ck[-1] + 4 * st.ncell <= st.ckpt.ptr + st.ckpt.nbytes
        assert recomputed == []
    else:
        # separate Born pass: the background of each segment, then its Born steps read it
        for n in range(1, nt + 1):
            if (n - 1) % k == 0:
                st.background_segment((n - 1) // k)
            for kind in range(nf):
                assert st.get(kind, n, None) in seg
        assert recomputed == [(j, True) for j in range(nseg)]
        recomputed.clear()
    # adjoint order: the last segment is at hand, every other one recomputed once
    for n in range(nt, 0, -1):
        for kind in range(nf):
            st.get(kind, n, None)
    assert recomputed == [(j, False) for j in range(nseg - 2, -1, -1)]


def test_memory_formula():
    assert born.checkpoint_fields(200, 1, 20) == 2 * 10 + 20 + 6
    assert born.checkpoint_fields(3000, 2, 77) == 2 * 2 * 39 + 77 * 2 + 12
    assert born.checkpoint_fields(10, 1, 50) == 2 + 10 + 6


@pytest.mark.parametrize("nt,k", [(37, 8), (48, 48), (200, 48), (10, 16), (1, 4)])
@pytest.mark.parametrize("nf", [1, 2])
def test_predictive_store_layout(monkeypatch, nt, k, nf):
    # every residual code, keyframe and segment slot of the predictive store is distinct and
    # inside its buffer; keyframes are the first frame of each segment
    monkeypatch.setattr(born.ginsu, "DeviceBuffer", _Buf)
    shape = (20, 18, 17)
    st = born._PredictiveStore(shape, nt, nf, k, 3, 16)
    codes, keys, slots = set(), set(), set()
    for n in range(1, nt + 1):
        for kind in range(nf):
            if (n - 1) % k == 0:
                keys.add(st._key(kind, (n - 1) // k))
            else:
                c = st._code(kind, n)
                assert st.codes.ptr <= c and c + 4 * st.words <= st.codes.ptr + st.codes.nbytes
                codes.add(c)
            if (n - 1) // k == 0:
                s = st._slot(kind, n)
                assert st.seg.ptr <= s and s + 4 * st.ncell <= st.seg.ptr + st.seg.nbytes
                slots.add(s)
    nseg = (nt + k - 1) // k
    assert len(keys) == nseg * nf and len(codes) == (nt - nseg) * nf
    assert len(slots) == min(k, nt) * nf
    assert max(keys) + 4 * st.kwords <= st.keys.ptr + st.keys.nbytes
    assert st.stored_bytes() == 4 * (st.kwords * nseg + st.words * (nt - nseg)) * nf
