"""ctypes loading of the in-tree HIP libraries with the reference's C signatures.

The product path has no CPU fallback: if a library is missing, loading raises.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(HERE, "lib")

_L, _F, _S = ctypes.c_long, ctypes.c_float, ctypes.c_size_t
_P = ctypes.c_void_p
_FP = ctypes.c_void_p  # float* passed as an integer address (host or device)

PREFIXES = {
    "iso": "Prop3DAcoIsoDenQ_DEO2_FDTD",
    "vti": "Prop3DAcoVTIDenQ_DEO2_FDTD",
    "tti": "Prop3DAcoTTIDenQ_DEO2_FDTD",
    "iso2d": "Prop2DAcoIsoDenQ_DEO2_FDTD",
    "vti2d": "Prop2DAcoVTIDenQ_DEO2_FDTD",
    "tti2d": "Prop2DAcoTTIDenQ_DEO2_FDTD",
}

# name -> (restype, argtypes); common to the three propagator libraries
_COMMON = {
    "alloc": (_P, [_L, _L, _L, _L, _L, _L, _F, _F, _F, _F, _L, _L, _L]),
    "free": (None, [_P]),
    "SetupDtOmegaInvQ": (None, [_P, _F, _F, _F]),
    "TimeStep": (None, [_P]),
    "ScaleSpatialDerivatives": (None, [_P]),
    "getNx": (_L, [_P]),
    "getNy": (_L, [_P]),
    "getNz": (_L, [_P]),
    # additive device path
    "SetHostCoherence": (None, [_P, _L]),
    "GetHostCoherence": (_L, [_P]),
    "SyncToDevice": (None, [_P]),
    "SyncToHost": (None, [_P]),
    "Synchronize": (None, [_P]),
    "GetDevicePtr": (_P, [_P, ctypes.c_char_p]),
    "GetStream": (_P, [_P]),
    "SetSources": (_L, [_P, _P, _S, _P, _P, _S, _L]),
    "SetReceivers": (_L, [_P, _P, _S, _P, _S, _L]),
    "SetSourcesInterp": (_L, [_P, _P, _S, _P, _P, _S, _P, _S, _S, _L]),
    "SetReceiversInterp": (_L, [_P, _P, _S, _P, _S, _P, _S, _S, _L]),
    "InjectRegistered": (_L, [_P, _L]),
    "ExtractRegistered": (_L, [_P, _L]),
    "TimeStepInjectExtract": (None, [_P, _L]),
    "GetReceiverData": (_L, [_P, _P]),
    "SetFusedStep": (None, [_P, _L]),
    "SetFusedTile": (_L, [_P, _L]),
    "GetFusedTile": (_L, [_P]),
    "TimeStepInjectExtractGraph": (_L, [_P, _L, _L]),
    "SetFusedConfig": (_L, [_P, _L, _L, _L]),
    "GetFusedConfig": (_L, [_P, _P]),
    "TimeStepBatch": (_L, [_P, _L, _L, _P, _P, _P]),
    "LastError": (ctypes.c_char_p, []),
    "ClearError": (None, []),
}

_PHYS = {
    "iso": {
        "ForwardBornInjection_V": (None, [_P, _FP, _FP]),
        "ForwardBornInjection_VB": (None, [_P, _FP, _FP, _FP, _FP]),
        "ForwardBornInjection_B": (None, [_P, _FP, _FP, _FP]),
        "AdjointBornAccumulation_V": (None, [_P, _FP, _FP]),
        "AdjointBornAccumulation_VB": (None, [_P, _FP, _FP, _FP, _FP]),
        "AdjointBornAccumulation_B": (None, [_P, _FP, _FP, _FP]),
        "AdjointBornAccumulation_wavefieldsep": (None, [_P, _FP, _FP, _L]),
        "TimeStepSaveDP": (_L, [_P, _FP]),
        "TimeStepBornInjection_V": (_L, [_P, _FP, _FP]),
        "TimeStepBornCoupled_V": (_L, [_P, _FP, _FP]),
        "TimeStepAdjointBornAccumulation_V": (_L, [_P, _FP, _FP]),
        "getV": (_P, [_P]), "getB": (_P, [_P]), "getDtOmegaInvQ": (_P, [_P]),
        "getPSpace": (_P, [_P]), "getPCur": (_P, [_P]), "getPOld": (_P, [_P]),
    },
    "vti": {
        "TimeStepLinear": (None, [_P]),
        "ForwardBornInjection_V": (None, [_P, _FP, _FP, _FP]),
        "ForwardBornInjection_VEA": (None, [_P, _FP, _FP, _FP, _FP, _FP, _FP, _FP]),
        "AdjointBornAccumulation_V": (None, [_P, _FP, _FP, _FP]),
        "AdjointBornAccumulation_wavefieldsep_V": (None, [_P, _FP, _FP, _FP, _L]),
        "AdjointBornAccumulation_VEA": (None, [_P, _FP, _FP, _FP, _FP, _FP, _FP, _FP]),
        "TimeStepSaveDP": (_L, [_P, _FP, _FP]),
        "TimeStepBornInjection_V": (_L, [_P, _FP, _FP, _FP]),
        "TimeStepAdjointBornAccumulation_V": (_L, [_P, _FP, _FP, _FP]),
        "getV": (_P, [_P]), "getEps": (_P, [_P]), "getEta": (_P, [_P]), "getB": (_P, [_P]),
        "getF": (_P, [_P]), "getPSpace": (_P, [_P]), "getMSpace": (_P, [_P]), "getPCur": (_P, [_P]),
        "getPOld": (_P, [_P]), "getMCur": (_P, [_P]), "getMOld": (_P, [_P]), "getDtOmegaInvQ": (_P, [_P]),
    },
}
_PHYS["tti"] = {k: v for k, v in _PHYS["vti"].items() if k != "TimeStepLinear"}
for _a in ("SinTheta", "CosTheta", "SinPhi", "CosPhi"):
    _PHYS["tti"]["get" + _a] = (_P, [_P])

SPACETIME = {
    "injectdata_float": (None, [_P, _P, _S, _S, _P, _S, _P, _S]),
    "injectdata_float_noalloc": (None, [_P, _P, _S, _S, _P, _S, _P, _S]),
    "injectdata_double": (None, [_P, _P, _S, _S, _P, _S, _P, _S]),
    "extractdata_float": (None, [_P, _P, _P, _S, _P, _S, _S, _S]),
    "extractdata_double": (None, [_P, _P, _P, _S, _P, _S, _S, _S]),
    "interpadjoint_1d_float": (None, [_P, _P, _P, _S, _S, _S]),
    "interpadjoint_1d_double": (None, [_P, _P, _P, _S, _S, _S]),
    "interpforward_1d_float": (None, [_P, _P, _P, _S, _S, _S]),
    "interpforward_1d_double": (None, [_P, _P, _P, _S, _S, _S]),
    "interpadjoint_nd_float": (None, [_P, _P, _P, _S, _S, _S, _S, _S]),
    "interpadjoint_nd_double": (None, [_P, _P, _P, _S, _S, _S, _S, _S]),
    "interpforward_nd_float": (None, [_P, _P, _P, _S, _S, _S, _S, _S]),
    "interpforward_nd_double": (None, [_P, _P, _P, _S, _S, _S, _S, _S]),
    "spacetime_LastError": (ctypes.c_char_p, []),
}


CODEC = {
    "wavefd_codec_words": (_L, [_P, _P, _L]),
    "wavefd_codec_compress": (_L, [_P, _P, _P, _L, _P]),
    "wavefd_codec_decompress": (_L, [_P, _L, _P]),
    "wavefd_codec_LastError": (ctypes.c_char_p, []),
}


GINSU = {
    "wavefd_ginsu_extract": (_L, [_P, _L, _L, _L, _L, _L, _L, _P, _L, _L, _L, _P]),
    "wavefd_ginsu_accumulate": (_L, [_P, _L, _L, _L, _L, _L, _L, _P, _L, _L, _L, _L, _L, _L, _L, _L, _L, _P]),
    "wavefd_device_alloc": (_P, [_L]),
    "wavefd_device_free": (None, [_P]),
    "wavefd_device_copy": (_L, [_P, _P, _L, _P]),
    "wavefd_device_zero": (_L, [_P, _L, _P]),
    "wavefd_device_sync": (_L, [_P]),
    "wavefd_ginsu_LastError": (ctypes.c_char_p, []),
}


LCODEC = {
    "wavefd_lcodec_code_words": (_L, [_L, _L, _L]),
    "wavefd_lcodec_encode": (_L, [_P, _L, _L, _L, _P, _P]),
    "wavefd_lcodec_dequantize": (_L, [_P, _P, _P]),
    "wavefd_lcodec_dequantize_n": (_L, [_P, _L, _L, _L, _P, _P]),
    "wavefd_lcodec_code_words_l": (_L, [_L, _L, _L, _L]),
    "wavefd_lcodec_encode_l": (_L, [_P, _L, _L, _L, _L, _P, _P]),
    "wavefd_lcodec_dequantize_nl": (_L, [_P, _L, _L, _L, _L, _P, _P]),
    "wavefd_lcodec_nparams": (_L, []),
    "wavefd_lcodec_net_create": (_P, []),
    "wavefd_lcodec_net_free": (None, [_P]),
    "wavefd_lcodec_net_get": (_L, [_P, _P]),
    "wavefd_lcodec_net_set": (_L, [_P, _P]),
    "wavefd_lcodec_net_init": (_L, [_P, ctypes.c_ulong]),
    "wavefd_lcodec_forward": (_L, [_P, _P, _L, _L, _L, _L, ctypes.c_float, _P, _P, _P]),
    "wavefd_lcodec_grad": (_L, [_P, _P, _P, _L, _L, _L, _L, ctypes.c_float, _P, _P, _P, _P, _P]),
    "wavefd_lcodec_halfstep": (_L, [_P, _P, _P]),
    "wavefd_lcodec_encode_learned": (_L, [_P, _P, _L, _L, _L, _P, _P, _P, _P]),
    "wavefd_lcodec_encode_learned_l": (_L, [_P, _P, _L, _L, _L, _L, _P, _P, _P, _P]),
    "wavefd_lcodec_adam": (_L, [_P, ctypes.c_float, _P]),
    "wavefd_lcodec_decode": (_L, [_P, _P, _P, _P, _P]),
    "wavefd_lcodec_gather": (_L, [_P, _L, _L, _L, _P, _L, _L, ctypes.c_float, _P, _P]),
    "wavefd_lcodec_norm": (ctypes.c_float, [_P]),
    "wavefd_lcodec_combine": (_L, [_L, _F, _P, _F, _P, _F, _P, _P, _P]),
    "wavefd_lcodec_predict_residual": (_L, [_L, _F, _P, _F, _P, _F, _P, _P, _P, _P, _P]),
    "wavefd_lcodec_LastError": (ctypes.c_char_p, []),
}

# the 2D Iso library (src/prop2DAcoIsoDenQ_DEO2_FDTD.cc), names without the prefix
_ISO2D = {
    "alloc": (_P, [_L, _L, _L, _L, _L, _F, _F, _F, _L, _L]),
    "free": (None, [_P]),
    "SetupDtOmegaInvQ": (None, [_P, _F, _F, _F]),
    "TimeStep": (None, [_P]),
    "ScaleSpatialDerivatives": (None, [_P]),
    "ForwardBornInjection_V": (None, [_P, _FP, _FP]),
    "ForwardBornInjection_VB": (None, [_P, _FP, _FP, _FP, _FP]),
    "ForwardBornInjection_B": (None, [_P, _FP, _FP, _FP]),
    "AdjointBornAccumulation_V": (None, [_P, _FP, _FP]),
    "AdjointBornAccumulation_VB": (None, [_P, _FP, _FP, _FP, _FP]),
    "AdjointBornAccumulation_B": (None, [_P, _FP, _FP, _FP]),
    "AdjointBornAccumulation_wavefieldsep": (None, [_P, _FP, _FP, _L]),
    "getNx": (_L, [_P]),
    "getNz": (_L, [_P]),
    "getV": (_P, [_P]),
    "getB": (_P, [_P]),
    "getDtOmegaInvQ": (_P, [_P]),
    "getPSpace": (_P, [_P]),
    "getPCur": (_P, [_P]),
    "getPOld": (_P, [_P]),
    "SetHostCoherence": (None, [_P, _L]),
    "SyncToDevice": (None, [_P]),
    "SyncToHost": (None, [_P]),
    "GetDevicePtr": (_P, [_P, ctypes.c_char_p]),
    "GetStream": (_P, [_P]),
    "Synchronize": (None, [_P]),
    "SetFusedStep": (None, [_P, _L]),
    "TimeStepSaveDP": (_L, [_P, _FP]),
    "TimeStepBornInjection_V": (_L, [_P, _FP, _FP]),
    "TimeStepAdjointBornAccumulation_V": (_L, [_P, _FP, _FP]),
    "LastError": (ctypes.c_char_p, []),
    "ClearError": (None, []),
}

# the 2D VTI and TTI libraries (src/prop2DAcoVTIDenQ_DEO2_FDTD.cc, src/prop2DAcoTTIDenQ_DEO2_FDTD.cc)
_ANISO2D = {
    "alloc": (_P, [_L, _L, _L, _L, _L, _F, _F, _F, _L, _L]),
    "free": (None, [_P]),
    "SetupDtOmegaInvQ": (None, [_P, _F, _F, _F]),
    "TimeStep": (None, [_P]),
    "ScaleSpatialDerivatives": (None, [_P]),
    "ForwardBornInjection_V": (None, [_P, _FP, _FP, _FP]),
    "ForwardBornInjection_VEA": (None, [_P, _FP, _FP, _FP, _FP, _FP, _FP, _FP]),
    "AdjointBornAccumulation_V": (None, [_P, _FP, _FP, _FP]),
    "AdjointBornAccumulation_wavefieldsep_V": (None, [_P, _FP, _FP, _FP, _L]),
    "AdjointBornAccumulation_VEA": (None, [_P, _FP, _FP, _FP, _FP, _FP, _FP, _FP]),
    "getNx": (_L, [_P]),
    "getNz": (_L, [_P]),
    "getV": (_P, [_P]),
    "getEps": (_P, [_P]),
    "getEta": (_P, [_P]),
    "getB": (_P, [_P]),
    "getF": (_P, [_P]),
    "getDtOmegaInvQ": (_P, [_P]),
    "getPSpace": (_P, [_P]),
    "getMSpace": (_P, [_P]),
    "getPCur": (_P, [_P]),
    "getPOld": (_P, [_P]),
    "getMCur": (_P, [_P]),
    "getMOld": (_P, [_P]),
    "SetHostCoherence": (None, [_P, _L]),
    "SyncToDevice": (None, [_P]),
    "SyncToHost": (None, [_P]),
    "GetDevicePtr": (_P, [_P, ctypes.c_char_p]),
    "GetStream": (_P, [_P]),
    "Synchronize": (None, [_P]),
    "SetFusedStep": (None, [_P, _L]),
    "TimeStepSaveDP": (_L, [_P, _FP, _FP]),
    "TimeStepBornInjection_V": (_L, [_P, _FP, _FP, _FP]),
    "TimeStepAdjointBornAccumulation_V": (_L, [_P, _FP, _FP, _FP]),
    "LastError": (ctypes.c_char_p, []),
    "ClearError": (None, []),
}
_VTI2D = dict(_ANISO2D, TimeStepLinear=(None, [_P]),
              applyFirstDerivatives2D_PlusHalf=(None, [_L, _L, _L, _L, _F, _F, _F, _F, _F, _F, _FP, _FP, _FP, _FP,
                                                       _L, _L]))
_TTI2D = dict(_ANISO2D, getSinTheta=(_P, [_P]), getCosTheta=(_P, [_P]))
_2D = {"iso2d": _ISO2D, "vti2d": _VTI2D, "tti2d": _TTI2D}

_TABLES = {"spacetime": SPACETIME, "codec": CODEC, "ginsu": GINSU, "lcodec": LCODEC}
_LIBNAMES = {"spacetime": "libspacetime.so", "codec": "libwavefd_codec.so", "ginsu": "libwavefd_ginsu.so",
             "lcodec": "libwavefd_lcodec.so"}
_ERRFN = {"spacetime": "spacetime_LastError", "codec": "wavefd_codec_LastError",
          "ginsu": "wavefd_ginsu_LastError", "lcodec": "wavefd_lcodec_LastError"}


def signatures(kind):
    """{symbol: (restype, argtypes)} for 'iso' | 'vti' | 'tti' | 'spacetime' | 'codec' | 'ginsu' |
    'lcodec'."""
    if kind in _TABLES:
        return dict(_TABLES[kind])
    pre = PREFIXES[kind]
    if kind in _2D:
        return {pre + "_" + name: sig for name, sig in _2D[kind].items()}
    out = {}
    for name, sig in list(_COMMON.items()) + list(_PHYS[kind].items()):
        out[pre + "_" + name] = sig
    return out


def lib_path(kind):
    if kind in _LIBNAMES:
        name = _LIBNAMES[kind]
    else:  # the reference's library names start in lower case: libprop3DAco...
        pre = PREFIXES[kind]
        name = "lib" + pre[0].lower() + pre[1:] + ".so"
    return os.path.join(LIBDIR, name)


_cache = {}


class _Bound:
    """Attribute access to a library's symbols without the common prefix."""

    def __init__(self, kind, dll):
        self.kind = kind
        self.dll = dll
        self.prefix = PREFIXES[kind] + "_" if kind in PREFIXES else ""

    def __getattr__(self, name):
        fn = getattr(self.dll, self.prefix + name)
        setattr(self, name, fn)
        return fn


def load(kind):
    """Load (once) and type one of the HIP libraries; raises when it has not been built."""
    if kind in _cache:
        return _cache[kind]
    path = lib_path(kind)
    if not os.path.exists(path):
        raise RuntimeError("%s is not built; run `python -c 'import __graft_entry__ as g; g.build()'` "
                           "or `python wavefd_amd/build.py`" % path)
    dll = ctypes.CDLL(path)
    for sym, (res, args) in signatures(kind).items():
        fn = getattr(dll, sym)
        fn.restype = res
        fn.argtypes = args
    b = _Bound(kind, dll)
    _cache[kind] = b
    return b


def last_error(kind):
    lib = load(kind)
    msg = getattr(lib, _ERRFN[kind])() if kind in _ERRFN else lib.LastError()
    return msg.decode() if msg else ""
