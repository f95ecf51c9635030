"""ctypes binding of libtalsh_amd.so, the MI355X TAL-SH C-ABI library.

This module is plumbing for tests and benchmarks: every call goes through the
C entry points declared in include/talsh.h and include/talsh_amd.h, with the
reference's names and argument meanings.  It performs no computation itself
and has no fallback: if the HIP library is missing, loading fails loudly.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtalsh_amd.so")

# constants (include/tensor_algebra.h, include/talsh.h)
NO_TYPE, R4, R8, C4, C8 = 0, 4, 8, 14, 18
DEV_NULL = DEV_DEFAULT = -1
DEV_HOST, DEV_NVIDIA_GPU, DEV_INTEL_MIC, DEV_AMD_GPU = 0, 1, 2, 3
NOPE, YEP = 0, 1
MEM_ALLOC_REGULAR, MEM_ALLOC_TMP_BUF, MEM_ALLOC_ALL_BUF = 0, 1, 2
# fast-math split methods of the R4 GEMM (include/talsh_amd.h)
SPLIT_BF16X6, SPLIT_FP16X3 = 0, 1
TALSH_TENSOR_CONTRACT = 82
TALSH_OP_EMPTY, TALSH_OP_PARTIAL, TALSH_OP_DEFINED, TALSH_OP_RETIRED = 0, 1, 2, 8
TRY_LATER, DEVICE_UNABLE, NOT_CLEAN = -918273645, -546372819, -192837465
COPY_D, COPY_M, COPY_T, COPY_K = 0, 1, 2, 3
TALSH_SUCCESS = 0
TALSH_FAILURE = -666
TALSH_NOT_AVAILABLE = -888
TALSH_NOT_IMPLEMENTED = -999
TALSH_NOT_INITIALIZED = 1000000
TALSH_ALREADY_INITIALIZED = 1000001
TALSH_INVALID_ARGS = 1000002
TALSH_OBJECT_NOT_EMPTY = 1000004
TALSH_OBJECT_IS_EMPTY = 1000005
TALSH_IN_PROGRESS = 1000006
TALSH_NOT_FOUND = 1000009
TALSH_TASK_ERROR = 1999999
TALSH_TASK_EMPTY = 2000000
TALSH_TASK_SCHEDULED = 2000001
TALSH_TASK_COMPLETED = 2000005
NP_DTYPE = {R4: np.float32, R8: np.float64, C4: np.complex64, C8: np.complex128}


def coh(*letters):
    """Coherence code from letters, e.g. coh('M','T','T') == COPY_MTT."""
    v = 0
    for x in letters:
        v = (v << 2) | {"D": 0, "M": 1, "T": 2, "K": 3}[x]
    return v


class TalshTens(ctypes.Structure):
    _fields_ = [("shape_p", ctypes.c_void_p), ("dev_rsc", ctypes.c_void_p), ("data_kind", ctypes.POINTER(ctypes.c_int)),
                ("avail", ctypes.POINTER(ctypes.c_int)), ("dev_rsc_len", ctypes.c_int), ("ndev", ctypes.c_int)]


class TalshTensArg(ctypes.Structure):
    _fields_ = [("tens_p", ctypes.c_void_p), ("source_image", ctypes.c_int)]


class TalshTask(ctypes.Structure):
    _fields_ = [("task_p", ctypes.c_void_p), ("task_error", ctypes.c_int), ("dev_kind", ctypes.c_int),
                ("data_kind", ctypes.c_int), ("coherence", ctypes.c_int), ("num_args", ctypes.c_int),
                ("tens_args", TalshTensArg * 4), ("data_vol", ctypes.c_double), ("flops", ctypes.c_double),
                ("exec_time", ctypes.c_double)]


class TensSignature(ctypes.Structure):
    _fields_ = [("num_dim", ctypes.c_int), ("offsets", ctypes.POINTER(ctypes.c_size_t))]


class TensShape(ctypes.Structure):
    _fields_ = [("num_dim", ctypes.c_int), ("dims", ctypes.POINTER(ctypes.c_int)),
                ("divs", ctypes.POINTER(ctypes.c_int)), ("grps", ctypes.POINTER(ctypes.c_int))]


class TalshTensSlice(ctypes.Structure):
    _fields_ = [("tensor", ctypes.POINTER(TalshTens)), ("bases", TensSignature), ("shape", TensShape)]


class TalshTensOp(ctypes.Structure):
    _fields_ = [("opkind", ctypes.c_int), ("data_kind", ctypes.c_int), ("num_args", ctypes.c_uint),
                ("tens_slice", TalshTensSlice * 4), ("symb_pattern", ctypes.c_char_p),
                ("alpha_real", ctypes.c_double), ("alpha_imag", ctypes.c_double), ("tens_arg", TalshTens * 4),
                ("task_handle", TalshTask), ("exec_dev_id", ctypes.c_int), ("stage", ctypes.c_int),
                ("time_started", ctypes.c_double), ("time_scheduled", ctypes.c_double),
                ("time_completed", ctypes.c_double), ("time_finished", ctypes.c_double)]


class DevRsc(ctypes.Structure):
    _fields_ = [("dev_id", ctypes.c_int), ("gmem_p", ctypes.c_void_p), ("buf_entry", ctypes.c_int),
                ("mem_attached", ctypes.c_int)]


_I, _P, _D, _S = ctypes.c_int, ctypes.c_void_p, ctypes.c_double, ctypes.c_size_t
_TP, _KP = ctypes.POINTER(TalshTens), ctypes.POINTER(TalshTask)
_IP, _DP = ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)

# (restype, argtypes) of every exported C entry point used from Python.
SIGNATURES = {
    "talshValidDataKind": (_I, [_I, _IP]),
    "talshInit": (_I, [ctypes.POINTER(_S), _IP, _I, _IP, _I, _IP, _I, _IP]),
    "talshShutdown": (_I, []),
    "talshDeviceCount": (_I, [_I, _IP]),
    "talshFlatDevId": (_I, [_I, _I]),
    "talshKindDevId": (_I, [_I, _IP]),
    "talshDeviceState": (_I, [_I, _I]),
    "talshDeviceState_": (_I, [_I, _I]),
    "talshDeviceBusyLeast": (_I, [_I]),
    "talshDeviceMemorySize": (_S, [_I, _I]),
    "talshDeviceBufferSize": (_S, [_I, _I]),
    "talshDeviceTensorSize": (_S, [_I, _I]),
    "talshDeviceBufferFreeSize": (_S, [_I, _I]),
    "talshDeviceGetFlops": (_D, [_I, _I]),
    "talshStats": (_I, [_I, _I]),
    "talshTensorCreate": (_I, [ctypes.POINTER(_TP)]),
    "talshTensorClean": (_I, [_TP]),
    "talshTensorIsEmpty": (_I, [_TP]),
    "talshTensorConstruct": (_I, [_TP, _I, _I, _IP, _I, _P, _I, _P, _D, _D]),
    "talshTensorConstruct_": (_I, [_TP, _I, _I, _IP, _I, _P, _I, _P, _D, _D]),
    "talshTensorImportData": (_I, [_TP, _I, _P]),
    "talshTensorDestruct": (_I, [_TP]),
    "talshTensorDestroy": (_I, [_TP]),
    "talshTensorRank": (_I, [_TP]),
    "talshTensorVolume": (_S, [_TP]),
    "talshTensorDimExtents": (_IP, [_TP, _IP]),
    "talshTensorDataKind": (_I, [_TP, _IP, _IP]),
    "talshTensorInUse": (_I, [_TP]),
    "talshTensorPresence": (_I, [_TP, _IP, _IP, _IP, _I, _I]),
    "talshTensorGetBodyAccess": (_I, [_TP, ctypes.POINTER(_P), _I, _I, _I]),
    "talshTensorGetScalar": (_I, [_TP, _DP, _DP]),
    "talshTaskCreate": (_I, [ctypes.POINTER(_KP)]),
    "talshTaskClean": (_I, [_KP]),
    "talshTaskIsEmpty": (_I, [_KP]),
    "talshTaskDestruct": (_I, [_KP]),
    "talshTaskDestroy": (_I, [_KP]),
    "talshTaskDevId": (_I, [_KP, _IP]),
    "talshTaskStatus": (_I, [_KP]),
    "talshTaskComplete": (_I, [_KP, _IP, _IP]),
    "talshTaskWait": (_I, [_KP, _IP]),
    "talshTasksWait": (_I, [_I, _KP, _IP]),
    "talshTaskTime": (_I, [_KP, _DP, _DP, _DP, _DP, _DP]),
    "talshTensorPlace": (_I, [_TP, _I, _I, _P, _I, _KP]),
    "talshTensorDiscard": (_I, [_TP, _I, _I]),
    "talshTensorDiscardOther": (_I, [_TP, _I, _I]),
    "talshTensorInit": (_I, [_TP, _D, _D, _I, _I, _I, _KP]),
    "talshTensorScale": (_I, [_TP, _D, _D, _I, _I, _I, _KP]),
    "talshTensorCopy": (_I, [ctypes.c_char_p, _TP, _TP, _I, _I, _I, _KP]),
    "talshTensorAdd": (_I, [ctypes.c_char_p, _TP, _TP, _D, _D, _I, _I, _I, _KP]),
    "talshTensorContract": (_I, [ctypes.c_char_p, _TP, _TP, _TP, _D, _D, _I, _I, _I, _I, _KP]),
    "talshTensorContractXL": (_I, [ctypes.c_char_p, _TP, _TP, _TP, _D, _D, _I, _I, _I]),
    "talshSetMemAllocPolicyHost": (None, [_I, _I, ctypes.POINTER(_I)]),
    "talshTensorContractXL_": (_I, [ctypes.c_char_p, _TP, _TP, _TP, _D, _D, _I, _I, _I]),
    "talshTensorNorm2": (_I, [_TP, _TP, _I, _I, _I, _KP]),
    "talshTensorSlice": (_I, [_TP, _TP, _IP, _I, _I, _I, _I, _KP]),
    "talshTensorInsert": (_I, [_TP, _TP, _IP, _I, _I, _I, _I, _KP]),
    "talshDetermineOptimalDevice": (_I, [_TP, _TP, _TP]),
    "talshDeviceBufferBasePtr": (_P, [_I, _I]),
    "get_contr_pattern_sym": (None, [_IP, _IP, _IP, _IP, ctypes.c_char_p, _IP, _IP]),
    "talshTensorImageNorm1_cpu": (_D, [_TP]),
    "talshEnableFastMath": (_I, [_I, _I]),
    "talshDisableFastMath": (_I, [_I, _I]),
    "talshQueryFastMath": (_I, [_I, _I]),
    "talshAmdSetDeviceBufferSize": (_I, [_S]),
    "talshAmdSetNumStreams": (_I, [_I]),
    "talshAmdSetContractXLMemory": (_I, [_S]),
    "talshAmdSetFastMathSplit": (_I, [_I]),
    "talshAmdGetFastMathSplit": (_I, []),
    "talshAmdDeviceSynchronize": (_I, [_I]),
    "talshAmdBuildInfo": (_I, [ctypes.c_char_p, _I]),
    "talshAmdTensorNorm2NoAlloc": (_I, [_I, _P, _S, _P, _I, _P]),
    "talshAmdBatchCreate": (_I, [ctypes.POINTER(_P)]),
    "talshAmdBatchAdd": (_I, [_P, ctypes.c_char_p, _TP, _TP, _TP, _D, _D, _I]),
    "talshAmdBatchSize": (_I, [_P]),
    "talshAmdBatchExecute": (_I, [_P, _I, _I, _KP]),
    "talshAmdBatchStats": (_I, [_P, _DP, ctypes.POINTER(ctypes.c_longlong), _IP, _DP, _IP]),
    "talshAmdBatchDestroy": (_I, [_P]),
    "talshAmdServerCreate": (_I, [ctypes.POINTER(_P), _I, _S, _I]),
    "talshAmdServerSubmit": (_I, [_P, _P]),
    "talshAmdServerProgress": (_I, [_P, _I, ctypes.POINTER(ctypes.c_longlong), _DP, _IP, _IP]),
    "talshAmdServerPending": (_I, [_P, _IP, _IP]),
    "talshAmdServerStats": (_I, [_P, _DP, _I]),
    "talshAmdServerDestroy": (_I, [_P]),
}

_lib = None


def load(path=LIB_PATH):
    """Loads libtalsh_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise RuntimeError("libtalsh_amd.so not built: run __graft_entry__.build() (%s missing)" % path)
        lib = ctypes.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib


def check(err, what=""):
    if err != TALSH_SUCCESS:
        raise RuntimeError("%s failed with TAL-SH status %d" % (what, err))
    return err


def init(host_buf_size=1 << 28, gpus=(0,), dev_buffer_size=None, nstreams=None):
    lib = load()
    if dev_buffer_size is not None:
        check(lib.talshAmdSetDeviceBufferSize(dev_buffer_size), "talshAmdSetDeviceBufferSize")
    if nstreams is not None:
        check(lib.talshAmdSetNumStreams(nstreams), "talshAmdSetNumStreams")
    hb = ctypes.c_size_t(host_buf_size)
    ham = ctypes.c_int(0)
    gl = (ctypes.c_int * max(1, len(gpus)))(*gpus)
    return lib.talshInit(ctypes.byref(hb), ctypes.byref(ham), len(gpus), gl, 0, None, 0, None)


def shutdown():
    return load().talshShutdown()


# talsh_tens_init_i that leaves the body as allocated
_INIT_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p)
_NOINIT = _INIT_FN(lambda data, shape, signature: 0)


def gpu_flat(n=0):
    return load().talshFlatDevId(DEV_NVIDIA_GPU, n)


class Tensor:
    """Owns one talsh_tens_t (caller-allocated struct, as in the reference tests)."""

    def __init__(self, dk, dims, dev_id=0, init_val=0.0, data=None, ext=None, noinit=False):
        """ext: external body (numpy array or raw address); noinit: leave a new Host body uninitialised."""
        self.lib = load()
        self.t = TalshTens()
        self.dk = dk
        self.dims = list(dims)
        check(self.lib.talshTensorClean(ctypes.byref(self.t)), "talshTensorClean")
        d = (ctypes.c_int * max(1, len(dims)))(*dims)
        iv = complex(init_val)
        addr = ext if isinstance(ext, int) else (ext.ctypes.data if ext is not None else None)
        init = ctypes.cast(_NOINIT, ctypes.c_void_p) if noinit else None
        err = self.lib.talshTensorConstruct(ctypes.byref(self.t), dk, len(dims), d, dev_id, addr, -1, init,
                                            iv.real, iv.imag)
        self._ext = ext
        if err != 0:
            raise RuntimeError("talshTensorConstruct failed: %d" % err)
        if data is not None:
            self.upload(data)

    @property
    def ptr(self):
        return ctypes.byref(self.t)

    def volume(self):
        return int(np.prod(self.dims)) if self.dims else 1

    def upload(self, data):
        """Writes `data` into the Host image (talshTensorImportData)."""
        a = np.ascontiguousarray(np.asarray(data).reshape(-1), dtype=NP_DTYPE[self.dk])
        assert a.size == self.volume()
        check(self.lib.talshTensorImportData(self.ptr, self.dk, a.ctypes.data), "talshTensorImportData")

    def host_address(self):
        """Address of the Host image body (talshTensorGetBodyAccess)."""
        body = ctypes.c_void_p()
        check(self.lib.talshTensorGetBodyAccess(self.ptr, ctypes.byref(body), self.dk, 0, DEV_HOST),
              "talshTensorGetBodyAccess")
        return body.value

    def host_array(self):
        """Copy of the Host image (moves the tensor to Host first if needed)."""
        body = ctypes.c_void_p()
        err = self.lib.talshTensorGetBodyAccess(self.ptr, ctypes.byref(body), self.dk, 0, DEV_HOST)
        if err != 0:
            check(self.lib.talshTensorPlace(self.ptr, 0, DEV_HOST, None, COPY_M, None), "talshTensorPlace")
            check(self.lib.talshTensorGetBodyAccess(self.ptr, ctypes.byref(body), self.dk, 0, DEV_HOST),
                  "talshTensorGetBodyAccess")
        n = self.volume()
        dt = np.dtype(NP_DTYPE[self.dk])
        buf = (ctypes.c_char * (n * dt.itemsize)).from_address(body.value)
        return np.frombuffer(buf, dtype=dt).copy()

    def place(self, dev_id, copy_ctrl=COPY_M):
        check(self.lib.talshTensorPlace(self.ptr, dev_id, DEV_NULL, None, copy_ctrl, None), "talshTensorPlace")

    def images(self):
        n = ctypes.c_int()
        c = (ctypes.c_int * 16)()
        k = (ctypes.c_int * 16)()
        check(self.lib.talshTensorPresence(self.ptr, ctypes.byref(n), c, k, DEV_NULL, -1), "talshTensorPresence")
        return list(c[:n.value])

    def scalar(self):
        re, im = ctypes.c_double(), ctypes.c_double()
        check(self.lib.talshTensorGetScalar(self.ptr, ctypes.byref(re), ctypes.byref(im)), "talshTensorGetScalar")
        return complex(re.value, im.value)

    def destruct(self):
        if self.t.shape_p:
            check(self.lib.talshTensorDestruct(self.ptr), "talshTensorDestruct")

    def __del__(self):
        try:
            if self.t.shape_p and _lib is not None:
                _lib.talshTensorDestruct(ctypes.byref(self.t))
        except Exception:
            pass


class Task:
    def __init__(self):
        self.lib = load()
        self.t = TalshTask()
        self.lib.talshTaskClean(ctypes.byref(self.t))

    @property
    def ptr(self):
        return ctypes.byref(self.t)

    def wait(self):
        st = ctypes.c_int()
        check(self.lib.talshTaskWait(self.ptr, ctypes.byref(st)), "talshTaskWait")
        return st.value

    def complete(self):
        st, ierr = ctypes.c_int(), ctypes.c_int()
        done = self.lib.talshTaskComplete(self.ptr, ctypes.byref(st), ctypes.byref(ierr))
        return done, st.value, ierr.value

    def times(self):
        v = [ctypes.c_double() for _ in range(5)]
        check(self.lib.talshTaskTime(self.ptr, *[ctypes.byref(x) for x in v]), "talshTaskTime")
        return dict(zip(("total", "comput", "input", "output", "mmul"), [x.value for x in v]))

    def destruct(self):
        return self.lib.talshTaskDestruct(self.ptr)


class Batch:
    """talshAmdBatch handle: many contractions, one persistent launch, graph replays."""

    def __init__(self):
        self.lib = load()
        self.h = ctypes.c_void_p()
        check(self.lib.talshAmdBatchCreate(ctypes.byref(self.h)), "talshAmdBatchCreate")

    def add(self, pattern, d, l, r, alpha=1.0, accumulative=YEP):
        return self.lib.talshAmdBatchAdd(self.h, pattern.encode(), d.ptr, l.ptr, r.ptr, float(alpha), 0.0,
                                         accumulative)

    def execute(self, task=None, gpu=0):
        return self.lib.talshAmdBatchExecute(self.h, gpu, DEV_NVIDIA_GPU, task.ptr if task else None)

    def stats(self):
        f, w, c, t, e = ctypes.c_double(), ctypes.c_longlong(), ctypes.c_int(), ctypes.c_double(), ctypes.c_int()
        check(self.lib.talshAmdBatchStats(self.h, ctypes.byref(f), ctypes.byref(w), ctypes.byref(c),
                                          ctypes.byref(t), ctypes.byref(e)), "talshAmdBatchStats")
        return {"flops": f.value, "work_items": w.value, "tuned_classes": c.value, "tune_ms": t.value,
                "executions": e.value}

    def destroy(self):
        if self.h:
            check(self.lib.talshAmdBatchDestroy(self.h), "talshAmdBatchDestroy")
            self.h = ctypes.c_void_p()


class Request(ctypes.Structure):
    _fields_ = [("pattern", ctypes.c_char_p), ("dtens", _TP), ("ltens", _TP), ("rtens", _TP),
                ("scale_real", ctypes.c_double), ("scale_imag", ctypes.c_double), ("copy_ctrl", ctypes.c_int),
                ("accumulative", ctypes.c_int), ("tag", ctypes.c_longlong)]


SERVER_STATS = ("submitted", "completed", "failed", "deferred", "p50_latency_s", "p99_latency_s",
                "mean_latency_s", "p50_queue_wait_s", "p99_queue_wait_s", "tflops", "peak_transient_bytes",
                "budget_bytes")


class Server:
    """talshAmdServer handle: continuous admission of contraction requests."""

    def __init__(self, gpu=0, hbm_budget=200 << 30, max_inflight=32):
        self.lib = load()
        self.h = ctypes.c_void_p()
        check(self.lib.talshAmdServerCreate(ctypes.byref(self.h), gpu, hbm_budget, max_inflight), "ServerCreate")
        self._keep = {}
        n = 4096
        self._tags = (ctypes.c_longlong * n)()
        self._lat = (ctypes.c_double * n)()
        self._st = (ctypes.c_int * n)()

    def submit(self, tag, pattern, d, l, r, alpha=1.0, copy_ctrl=None, accumulative=YEP):
        a = complex(alpha)
        pb = pattern.encode()
        req = Request(pb, ctypes.pointer(d.t), ctypes.pointer(l.t), ctypes.pointer(r.t), a.real, a.imag,
                      coh("T", "T", "T") if copy_ctrl is None else copy_ctrl, accumulative, tag)
        err = self.lib.talshAmdServerSubmit(self.h, ctypes.byref(req))
        if err == 0:
            self._keep[tag] = (pb, d, l, r)
        return err

    def progress(self):
        n = ctypes.c_int()
        check(self.lib.talshAmdServerProgress(self.h, len(self._tags), self._tags, self._lat, self._st,
                                              ctypes.byref(n)), "ServerProgress")
        out = [(self._tags[i], self._lat[i], self._st[i]) for i in range(n.value)]
        for tag, _, _ in out:
            self._keep.pop(tag, None)
        return out

    def pending(self):
        q, f = ctypes.c_int(), ctypes.c_int()
        self.lib.talshAmdServerPending(self.h, ctypes.byref(q), ctypes.byref(f))
        return q.value, f.value

    def stats(self):
        v = (ctypes.c_double * 12)()
        check(self.lib.talshAmdServerStats(self.h, v, 12), "ServerStats")
        return dict(zip(SERVER_STATS, list(v)))

    def destroy(self):
        if self.h:
            check(self.lib.talshAmdServerDestroy(self.h), "ServerDestroy")
            self.h = ctypes.c_void_p()
            self._keep.clear()


def contract(pattern, d, l, r, alpha=1.0, dev_id=DEV_DEFAULT, dev_kind=DEV_DEFAULT, copy_ctrl=None,
             accumulative=YEP, task=None):
    copy_ctrl = coh("M", "T", "T") if copy_ctrl is None else copy_ctrl
    a = complex(alpha)
    return load().talshTensorContract(pattern.encode(), d.ptr, l.ptr, r.ptr, a.real, a.imag, dev_id, dev_kind,
                                      copy_ctrl, accumulative, task.ptr if task else None)


def contract_xl(pattern, d, l, r, alpha=1.0, dev_id=DEV_DEFAULT, dev_kind=DEV_DEFAULT, accumulative=YEP):
    """talshTensorContractXL: blocking; d ends with one image, on the Host."""
    a = complex(alpha)
    return load().talshTensorContractXL(pattern.encode(), d.ptr, l.ptr, r.ptr, a.real, a.imag, dev_id, dev_kind,
                                        accumulative)


def fast_math(enable, split=SPLIT_BF16X6):
    """talshEnableFastMath / talshDisableFastMath on the GPUs, with the split method of the R4 GEMM
    (talshAmdSetFastMathSplit); returns the status of the call that failed, else 0."""
    lib = load()
    if not enable:
        return lib.talshDisableFastMath(DEV_NVIDIA_GPU, DEV_DEFAULT)
    return lib.talshAmdSetFastMathSplit(split) or lib.talshEnableFastMath(DEV_NVIDIA_GPU, DEV_DEFAULT)


def set_xl_memory(nbytes):
    """talshAmdSetContractXLMemory: device bytes an XL contraction may use (0 = from the free pool)."""
    return load().talshAmdSetContractXLMemory(nbytes)


def slice(d, l, offsets, dev_id=DEV_DEFAULT, dev_kind=DEV_DEFAULT, copy_ctrl=None, accumulative=NOPE, task=None):
    """talshTensorSlice: d (+)= block of l at 0-based offsets."""
    copy_ctrl = coh("M", "T") if copy_ctrl is None else copy_ctrl
    off = (ctypes.c_int * max(1, len(offsets)))(*offsets)
    return load().talshTensorSlice(d.ptr, l.ptr, off, dev_id, dev_kind, copy_ctrl, accumulative,
                                   task.ptr if task else None)


def insert(d, l, offsets, dev_id=DEV_DEFAULT, dev_kind=DEV_DEFAULT, copy_ctrl=None, accumulative=NOPE, task=None):
    """talshTensorInsert: block of d at 0-based offsets (+)= l."""
    copy_ctrl = coh("M", "T") if copy_ctrl is None else copy_ctrl
    off = (ctypes.c_int * max(1, len(offsets)))(*offsets)
    return load().talshTensorInsert(d.ptr, l.ptr, off, dev_id, dev_kind, copy_ctrl, accumulative,
                                    task.ptr if task else None)


def pattern_sym(lrank, rrank, conj, dig):
    """get_contr_pattern_sym: digital -> symbolic contraction pattern; returns (ierr, string)."""
    buf = ctypes.create_string_buffer(4096)
    cpl, ierr = ctypes.c_int(), ctypes.c_int()
    lr, rr, cj = ctypes.c_int(lrank), ctypes.c_int(rrank), ctypes.c_int(conj)
    d = (ctypes.c_int * max(1, len(dig)))(*dig)
    load().get_contr_pattern_sym(ctypes.byref(lr), ctypes.byref(rr), ctypes.byref(cj), d, buf, ctypes.byref(cpl),
                                 ctypes.byref(ierr))
    return ierr.value, buf.raw[:cpl.value].decode() if ierr.value == 0 else ""


def copy(pattern, d, l, dev_id=DEV_DEFAULT, dev_kind=DEV_DEFAULT, copy_ctrl=None, task=None):
    copy_ctrl = coh("M", "T") if copy_ctrl is None else copy_ctrl
    return load().talshTensorCopy(pattern.encode(), d.ptr, l.ptr, dev_id, dev_kind, copy_ctrl,
                                  task.ptr if task else None)


def add(pattern, d, l, alpha=1.0, dev_id=DEV_DEFAULT, dev_kind=DEV_DEFAULT, copy_ctrl=None, task=None):
    copy_ctrl = coh("M", "T") if copy_ctrl is None else copy_ctrl
    a = complex(alpha)
    return load().talshTensorAdd(pattern.encode(), d.ptr, l.ptr, a.real, a.imag, dev_id, dev_kind, copy_ctrl,
                                 task.ptr if task else None)


def norm2(d, l, dev_id=DEV_DEFAULT, dev_kind=DEV_DEFAULT, copy_ctrl=None, task=None):
    copy_ctrl = coh("T", "T") if copy_ctrl is None else copy_ctrl
    return load().talshTensorNorm2(d.ptr, l.ptr, dev_id, dev_kind, copy_ctrl, task.ptr if task else None)


def scale(d, alpha, dev_id=DEV_DEFAULT, dev_kind=DEV_DEFAULT, copy_ctrl=COPY_M, task=None):
    a = complex(alpha)
    return load().talshTensorScale(d.ptr, a.real, a.imag, dev_id, dev_kind, copy_ctrl, task.ptr if task else None)


def init_value(d, val, dev_id=DEV_DEFAULT, dev_kind=DEV_DEFAULT, copy_ctrl=COPY_M, task=None):
    v = complex(val)
    return load().talshTensorInit(d.ptr, v.real, v.imag, dev_id, dev_kind, copy_ctrl, task.ptr if task else None)
