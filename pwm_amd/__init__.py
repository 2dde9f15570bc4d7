"""pwm_amd -- MI355X (gfx950) wavelet matrix / wavelet tree construction.

Host-side mirror of kurpicz/pwm's construction-algorithm interface
(include/benchmark/algorithm.hpp:39-191, include/wx_*.hpp) over the C-ABI
library libpwm_hip.so (include/pwm_hip.h).  Every computation runs in the HIP
library; if it is missing, importing this package raises.

    import pwm_amd
    for algo in pwm_amd.algorithm_list():          # algorithm_list<...>
        ws = algo.compute(text, len(text), levels)  # wx_*::compute
        ws.levels(), ws.bvs()[l], ws.zeros(), ws.level_bit_size(l)
"""
import ctypes
import os

import numpy as np

try:  # share one HIP runtime with torch: it must be loaded before libpwm_hip
    import torch  # noqa: F401
except ImportError:
    torch = None

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpwm_hip.so")
MAX_LEVELS = 64

OK = 0
ERR_INVALID = -1
ERR_NO_DEVICE = -2
ERR_OOM = -3
ERR_HIP = -4
ERR_SYMBOL_RANGE = -5
ERR_UNSUPPORTED = -6
ERR_BUDGET = -7
ERR_TIMEOUT = -8

FLAG_TREE = 1
FLAG_TEXT_ON_DEVICE = 2
FLAG_CHECK_RANGE = 4
FLAG_REUSE_OUTPUT = 8
FLAG_GRAPH = 16
FLAG_AUTOTUNE = 32
FLAG_STREAM = 64

# algorithm ids (pwm_hip.h enum pwmhip_algo)
ALGOS = {
    "pc": 0, "ps": 1, "ppc": 2, "pps": 3, "pc_ss": 4, "ppc_ss": 5, "dd_pc": 6, "dd_ps": 7,
    "dd_pc_ss": 8, "naive": 9, "huff_pc": 10, "huff_ps": 11, "huff_pc_ss": 12, "huff_naive": 13,
    "huff_dd_pc": 14,
}
# reference traits: is_parallel (wx_*.hpp `static constexpr bool is_parallel`)
_PARALLEL = {"ppc", "pps", "dd_pc", "dd_ps", "dd_pc_ss", "ppc_ss", "huff_dd_pc"}


class PwmError(RuntimeError):
    def __init__(self, status, what=""):
        self.status = status
        super().__init__("%s (status %d)%s" % (_status_string(status), status, (": " + what) if what else ""))


class _WS(ctypes.Structure):
    _fields_ = [
        ("levels", ctypes.c_uint64),
        ("text_size", ctypes.c_uint64),
        ("is_tree", ctypes.c_int32),
        ("is_huffman", ctypes.c_int32),
        ("d_words", ctypes.c_void_p),
        ("level_offset", ctypes.c_uint64 * MAX_LEVELS),
        ("level_bits", ctypes.c_uint64 * MAX_LEVELS),
        ("n_zeros", ctypes.c_uint64),
        ("zeros", ctypes.c_uint64 * MAX_LEVELS),
        ("n_codes", ctypes.c_uint64),
        ("code_len", ctypes.POINTER(ctypes.c_uint64)),
        ("code_word", ctypes.POINTER(ctypes.c_uint64)),
        ("alloc_bytes", ctypes.c_uint64),
    ]


class _ServerConfig(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("hbm_budget_bytes", ctypes.c_uint64), ("workers", ctypes.c_uint32),
                ("max_batch_texts", ctypes.c_uint32), ("max_batch_bytes", ctypes.c_uint64),
                ("max_query_batch", ctypes.c_uint64)]


class _ServerStats(ctypes.Structure):
    _fields_ = [(k, ctypes.c_uint64) for k in ("builds_done", "builds_failed", "queries_done",
                                               "query_requests_done", "build_launches", "query_launches")] + \
               [(k, ctypes.c_double) for k in ("build_p50_ms", "build_p99_ms", "query_p50_ms", "query_p99_ms")] + \
               [("bytes_in_use", ctypes.c_uint64), ("bytes_peak", ctypes.c_uint64),
                ("mean_texts_per_launch", ctypes.c_double)]


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError("pwm_amd: %s is missing -- run __graft_entry__.build()" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    vp, u64, u32 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32
    sig = {
        "pwmhip_create": [ctypes.POINTER(vp), ctypes.c_int],
        "pwmhip_destroy": [vp],
        "pwmhip_device_count": [ctypes.POINTER(ctypes.c_int)],
        "pwmhip_build": [vp, vp, u64, u32, u32, u32, u32, ctypes.POINTER(_WS)],
        "pwmhip_build_noalloc": [vp, vp, u64, u32, u32, u32, u32, vp, u64, ctypes.POINTER(_WS)],
        "pwmhip_ws_free": [vp, ctypes.POINTER(_WS)],
        "pwmhip_ws_download": [vp, ctypes.POINTER(_WS), u64, vp],
        "pwmhip_rs_build": [vp, ctypes.POINTER(_WS), ctypes.POINTER(vp)],
        "pwmhip_rs_free": [vp, vp],
        "pwmhip_bv_query": [vp, vp, vp, vp, vp, u64, vp, u32],
        "pwmhip_query": [vp, ctypes.POINTER(_WS), vp, vp, vp, vp, u64, vp, u32],
        "pwmhip_copy_bits": [vp, vp, vp, u64, u64, u64],
        "pwmhip_reduce_alphabet": [vp, vp, u64, u32, ctypes.POINTER(u64)],
        "pwmhip_build_batch": [vp, vp, vp, u64, u32, u32, u32, u32, ctypes.POINTER(_WS)],
        "pwmhip_last_timing": [vp, ctypes.POINTER(ctypes.c_double)],
        "pwmhip_generate": [u32, u64, u64, u32, u32, vp],
        "pwmhip_server_create": [ctypes.POINTER(_ServerConfig), ctypes.POINTER(vp)],
        "pwmhip_server_destroy": [vp],
        "pwmhip_server_build": [vp, vp, u64, u32, u32, u32, u32, ctypes.POINTER(u64)],
        "pwmhip_server_query": [vp, u64, vp, vp, vp, u64, vp, ctypes.POINTER(u64)],
        "pwmhip_server_wait": [vp, u64, ctypes.c_int64],
        "pwmhip_server_download": [vp, u64, u64, vp],
        "pwmhip_server_index_info": [vp, u64, ctypes.POINTER(_WS)],
        "pwmhip_server_release": [vp, u64],
        "pwmhip_server_stats_get": [vp, ctypes.POINTER(_ServerStats)],
    }
    for name, args in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = ctypes.c_int
    lib.pwmhip_status_string.argtypes = [ctypes.c_int]
    lib.pwmhip_status_string.restype = ctypes.c_char_p
    return lib


_lib = _load()


def lib():
    return _lib


def _status_string(st):
    return _lib.pwmhip_status_string(st).decode()


def _check(st, what=""):
    if st != OK:
        raise PwmError(st, what)


def device_count():
    c = ctypes.c_int(0)
    _lib.pwmhip_device_count(ctypes.byref(c))
    return c.value


def generate(kind, seed, size, word_width=1, levels=8):
    """Seeded synthetic text (host numpy array): kind 'uniform', 'dna' or 'zipf'."""
    k = {"uniform": 0, "dna": 1, "zipf": 2}[kind]
    dt = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}[word_width]
    out = np.empty(size, dt)
    _check(_lib.pwmhip_generate(k, seed, size, word_width, levels, out.ctypes.data_as(ctypes.c_void_p)),
           "generate")
    return out


class Context:
    """One HIP device, one stream, one scratch arena (pwmhip_ctx)."""

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        _check(_lib.pwmhip_create(ctypes.byref(h), device), "create")
        self.h = h
        self.device = device

    def close(self):
        if self.h:
            _lib.pwmhip_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def last_timing(self):
        arr = (ctypes.c_double * 4)()
        _check(_lib.pwmhip_last_timing(self.h, arr))
        return list(arr)


_default_ctx = None


def default_context():
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


def _text_arg(text):
    """(pointer, size, word_width, flags) for a numpy array or a device tensor."""
    if hasattr(text, "data_ptr") and hasattr(text, "is_cuda"):
        if not text.is_cuda:
            text = text.numpy()
        else:
            return ctypes.c_void_p(text.data_ptr()), text.numel(), text.element_size(), FLAG_TEXT_ON_DEVICE
    if isinstance(text, (bytes, bytearray, str)):
        text = np.frombuffer(text.encode("utf-8") if isinstance(text, str) else bytes(text), np.uint8)
    text = np.ascontiguousarray(text)
    if text.dtype not in (np.uint8, np.uint16, np.uint32, np.uint64):
        raise PwmError(ERR_INVALID, "text must hold unsigned 8/16/32/64-bit symbols")
    return text.ctypes.data_as(ctypes.c_void_p), text.size, text.itemsize, 0, text


class WaveletStructure:
    """Result of a construction (reference: construction/wavelet_structure.hpp:
    142-199).  Levels stay in device memory; bvs() downloads them on first use."""

    def __init__(self, ctx, ws):
        self._ctx = ctx
        self._ws = ws
        self._bvs = None

    # -- reference accessors ------------------------------------------------
    def levels(self):
        return int(self._ws.levels)

    def text_size(self):
        return int(self._ws.text_size)

    def is_tree(self):
        return bool(self._ws.is_tree)

    def is_huffman_shaped(self):
        return bool(self._ws.is_huffman)

    def level_bit_size(self, level):
        return int(self._ws.level_bits[level])

    def zeros(self):
        return np.array([self._ws.zeros[i] for i in range(self._ws.n_zeros)], dtype=np.uint64)

    def bvs(self):
        if self._bvs is None:
            out = []
            for l in range(self.levels()):
                nw = (self.level_bit_size(l) + 63) // 64
                a = np.zeros(max(1, nw), np.uint64)
                if nw:
                    _check(_lib.pwmhip_ws_download(self._ctx.h, ctypes.byref(self._ws), l,
                                                   a.ctypes.data_as(ctypes.c_void_p)), "download")
                out.append(a[:nw])
            self._bvs = out
        return self._bvs

    def codes(self):
        """Huffman code (length, word) per symbol 0..max_symbol."""
        n = int(self._ws.n_codes)
        if not n:
            return None
        ln = np.ctypeslib.as_array(self._ws.code_len, shape=(n,)).copy()
        wd = np.ctypeslib.as_array(self._ws.code_word, shape=(n,)).copy()
        return ln, wd

    # -- device side --------------------------------------------------------
    def device_words(self):
        return int(self._ws.d_words or 0)

    def level_offset(self, level):
        return int(self._ws.level_offset[level])

    def free(self):
        if self._ws is not None and self._ctx.h:
            _lib.pwmhip_ws_free(self._ctx.h, ctypes.byref(self._ws))
        self._ws = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def build(text, levels, algo="ppc", is_tree=False, ctx=None, check_range=False, size=None, reuse=None,
          stream=False):
    """pwmhip_build: construct one structure; `text` is a numpy array (host)
    or a device tensor (torch on cuda:0).  `reuse` is an earlier
    WaveletStructure whose device memory is taken over.  `stream` (host
    texts, 2..4 levels): the text is copied in slices while the local build
    runs and is never resident in HBM as a whole."""
    ctx = ctx or default_context()
    arg = _text_arg(text)
    ptr, n, w, flags = arg[0], arg[1], arg[2], arg[3]
    keep = arg[4] if len(arg) > 4 else text  # keep host array alive
    if size is not None:
        n = size
    if is_tree:
        flags |= FLAG_TREE
    if check_range:
        flags |= FLAG_CHECK_RANGE
    if stream:
        flags |= FLAG_STREAM
    ws = _WS()
    # a structure of a batch is a view into the batch's shared block, which
    # its siblings still use: only a structure that owns its memory is
    # taken over
    if reuse is not None and reuse._ws is not None and getattr(reuse, "_block", None) is None:
        ws = reuse._ws
        reuse._ws = None
        flags |= FLAG_REUSE_OUTPUT
    st = _lib.pwmhip_build(ctx.h, ptr, n, w, levels, ALGOS[algo], flags, ctypes.byref(ws))
    del keep
    _check(st, "build %s" % algo)
    return WaveletStructure(ctx, ws)


class _BatchBlock:
    """Device memory shared by the structures of one build_batch call; freed
    when the last of them is gone."""

    def __init__(self, ctx, ws):
        self._ctx = ctx
        self._ws = ws

    def __del__(self):
        try:
            if self._ws is not None and self._ctx.h:
                _lib.pwmhip_ws_free(self._ctx.h, ctypes.byref(self._ws))
        except Exception:
            pass
        self._ws = None


def build_batch(texts, levels, algo="ppc", is_tree=False, ctx=None, check_range=False, graph=False,
                autotune=False, reuse=None):
    """pwmhip_build_batch: construct one structure per text in one launch per
    phase.  `texts` is a list of numpy arrays (host) or of device tensors
    (all of one kind and symbol width).  Returns a list of WaveletStructure;
    each is equal to build(text, ...) of its text.  `graph` replays the
    captured launch sequence when a batch repeats the previous one's
    geometry and addresses; `autotune` times the launch variants once per
    (levels, mean length) bucket of the context.  `reuse` is the list an
    earlier build_batch returned: its device block is taken over (and those
    structures become invalid)."""
    ctx = ctx or default_context()
    T = len(texts)
    if T == 0:
        return []
    args = [_text_arg(t) for t in texts]
    ptrs = (ctypes.c_void_p * T)(*[a[0].value for a in args])
    sizes = (ctypes.c_uint64 * T)(*[a[1] for a in args])
    widths = {a[2] for a in args}
    kinds = {a[3] for a in args}
    if len(widths) != 1 or len(kinds) != 1:
        raise PwmError(ERR_INVALID, "build_batch: texts must share symbol width and location")
    keep = [a[4] if len(a) > 4 else t for a, t in zip(args, texts)]
    flags = kinds.pop()
    if is_tree:
        flags |= FLAG_TREE
    if check_range:
        flags |= FLAG_CHECK_RANGE
    if graph:
        flags |= FLAG_GRAPH
    if autotune:
        flags |= FLAG_AUTOTUNE
    outs = (_WS * T)()
    if reuse:
        blk = getattr(reuse[0], "_block", None)
        if blk is not None and blk._ws is not None:
            ctypes.memmove(ctypes.byref(outs[0]), ctypes.byref(blk._ws), ctypes.sizeof(_WS))
            blk._ws = None
            flags |= FLAG_REUSE_OUTPUT
        for r in reuse:
            r._ws = None
            r._block = None
    st = _lib.pwmhip_build_batch(ctx.h, ptrs, sizes, T, widths.pop(), levels, ALGOS[algo], flags, outs)
    del keep
    _check(st, "build_batch %s" % algo)
    res = []
    block = None
    if outs[0].alloc_bytes and (T == 1 or outs[1].alloc_bytes == 0):
        # one shared allocation, owned by the first structure: hand it to a
        # block object and make every structure a view that keeps it alive
        owner = _WS()
        ctypes.memmove(ctypes.byref(owner), ctypes.byref(outs[0]), ctypes.sizeof(_WS))
        owner.code_len = None
        owner.code_word = None
        owner.n_codes = 0
        outs[0].alloc_bytes = 0
        block = _BatchBlock(ctx, owner)
    for t in range(T):
        w = _WS()
        ctypes.memmove(ctypes.byref(w), ctypes.byref(outs[t]), ctypes.sizeof(_WS))
        ws = WaveletStructure(ctx, w)
        ws._block = block
        res.append(ws)
    return res


def reduce_alphabet(text, ctx=None):
    """pwmhip_reduce_alphabet: renumbers the symbols of a device tensor in
    place in first-occurrence order (reduce_alphabet,
    util/alphabet_util.hpp:18-45) and returns the number of levels
    (levels_for_max_char).  A numpy array is copied to the device, reduced
    there and returned as a new array: (array, levels)."""
    ctx = ctx or default_context()
    levels = ctypes.c_uint64(0)
    if hasattr(text, "data_ptr") and getattr(text, "is_cuda", False):
        _check(_lib.pwmhip_reduce_alphabet(ctx.h, ctypes.c_void_p(text.data_ptr()), text.numel(),
                                           text.element_size(), ctypes.byref(levels)), "reduce_alphabet")
        return int(levels.value)
    import torch
    arr = np.ascontiguousarray(text)
    signed = {1: np.int8, 2: np.int16, 4: np.int32, 8: np.int64}[arr.itemsize]
    d = torch.from_numpy(arr.view(signed).copy()).to("cuda:%d" % ctx.device)
    _check(_lib.pwmhip_reduce_alphabet(ctx.h, ctypes.c_void_p(d.data_ptr()), d.numel(), d.element_size(),
                                       ctypes.byref(levels)), "reduce_alphabet")
    torch.cuda.synchronize()
    return d.cpu().numpy().view(arr.dtype), int(levels.value)


class ConstructionAlgorithm:
    """Mirror of construction_algorithm (algorithm.hpp:83-191) for one GPU
    algorithm; `name` follows the reference's registrations
    (src/register/*.cpp), e.g. "wm_ppc" / "wt_pps" / "wm_huff_pc"."""

    def __init__(self, name, algo, is_tree, word_width, description):
        self._name = name
        self._algo = algo
        self._tree = is_tree
        self._width = word_width
        self._description = description

    def name(self):
        return self._name

    def description(self):
        return self._description

    def word_width(self):
        return self._width

    def is_tree(self):
        return self._tree

    def is_parallel(self):
        return self._algo in _PARALLEL

    def is_huffman_shaped(self):
        return self._algo.startswith("huff")

    def is_input_external(self):
        return False

    def is_output_external(self):
        return False

    def compute(self, text, size, levels, ctx=None):
        return build(text, levels, self._algo, self._tree, ctx=ctx, size=size)

    # compute_bitvector (algorithm.hpp:106-125) returns the structure
    compute_bitvector = compute


def algorithm_list(word_width=1):
    """Mirror of algorithm_list<false,false,w>::get_algorithm_list()."""
    out = []
    for algo in ALGOS:
        for tree in (False, True):
            name = ("wt_" if tree else "wm_") + algo
            out.append(ConstructionAlgorithm(name, algo, tree, word_width,
                                             "gfx950 HIP construction (%s)" % algo))
    return out


def get_algorithm(name, word_width=1):
    for a in algorithm_list(word_width):
        if a.name() == name:
            return a
    raise KeyError(name)


class QuerySupport:
    """Rank / select directories over every level plus batched queries
    (mirror of query_support<>, queries/query_support.hpp:18-186)."""

    ACCESS, RANK, SELECT = 0, 1, 2
    RANK1, RANK0, SELECT1, SELECT0 = 0, 1, 2, 3

    def __init__(self, ws):
        self._ws = ws
        self._ctx = ws._ctx
        h = ctypes.c_void_p()
        _check(_lib.pwmhip_rs_build(self._ctx.h, ctypes.byref(ws._ws), ctypes.byref(h)), "rs_build")
        self.h = h

    def free(self):
        if self.h and self._ctx.h:
            _lib.pwmhip_rs_free(self._ctx.h, self.h)
        self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def bv_query(self, level, op, arg):
        level = np.ascontiguousarray(level, np.uint32)
        op = np.ascontiguousarray(op, np.uint32)
        arg = np.ascontiguousarray(arg, np.uint64)
        res = np.zeros(max(1, op.size), np.uint64)
        vp = ctypes.c_void_p
        _check(_lib.pwmhip_bv_query(self._ctx.h, self.h, level.ctypes.data_as(vp), op.ctypes.data_as(vp),
                                    arg.ctypes.data_as(vp), op.size, res.ctypes.data_as(vp), 0), "bv_query")
        return res[:op.size]

    def query(self, op, a, b=None):
        """op 0 access(a); op 1 rank(symbol a, prefix b); op 2 select(symbol a, k b)."""
        op = np.ascontiguousarray(op, np.uint32)
        a = np.ascontiguousarray(a, np.uint64)
        b = np.ascontiguousarray(b if b is not None else np.zeros(op.size), np.uint64)
        res = np.zeros(max(1, op.size), np.uint64)
        vp = ctypes.c_void_p
        _check(_lib.pwmhip_query(self._ctx.h, ctypes.byref(self._ws._ws), self.h, op.ctypes.data_as(vp),
                                 a.ctypes.data_as(vp), b.ctypes.data_as(vp), op.size, res.ctypes.data_as(vp), 0),
               "query")
        return res[:op.size]

    def access(self, i):
        return int(self.query([0], [i])[0])

    def rank(self, symbol, index):
        return int(self.query([1], [symbol], [index])[0])

    def select(self, symbol, occ):
        return int(self.query([2], [symbol], [occ])[0])


class Server:
    """Serving loop on one device (pwmhip_server_*): builds and query batches
    are submitted and return at once; the native scheduler batches builds of
    one shape under the HBM budget, keeps indices resident, and answers the
    queries of all ready indices together.  Inputs stay referenced here until
    their request completes."""

    def __init__(self, device=0, hbm_budget_bytes=0, workers=0, max_batch_texts=0, max_batch_bytes=0,
                 max_query_batch=0):
        cfg = _ServerConfig(device, hbm_budget_bytes, workers, max_batch_texts, max_batch_bytes, max_query_batch)
        h = ctypes.c_void_p()
        _check(_lib.pwmhip_server_create(ctypes.byref(cfg), ctypes.byref(h)), "server_create")
        self.h = h
        self._keep = {}
        self._results = {}

    def close(self):
        if self.h:
            _lib.pwmhip_server_destroy(self.h)
            self.h = None
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def build(self, text, levels, algo="ppc", is_tree=False, check_range=False):
        """Submits a build of a host text; returns the index id."""
        text = np.ascontiguousarray(text)
        if text.dtype not in (np.uint8, np.uint16, np.uint32, np.uint64):
            raise PwmError(ERR_INVALID, "text must hold unsigned 8/16/32/64-bit symbols")
        flags = (FLAG_TREE if is_tree else 0) | (FLAG_CHECK_RANGE if check_range else 0)
        rid = ctypes.c_uint64()
        _check(_lib.pwmhip_server_build(self.h, text.ctypes.data_as(ctypes.c_void_p), text.size, text.itemsize,
                                        levels, ALGOS[algo], flags, ctypes.byref(rid)), "server_build")
        self._keep[rid.value] = (text,)
        return rid.value

    def query(self, index, op, a, b=None):
        """Submits symbol queries (op 0 access(a), 1 rank(a, b), 2 select(a, b));
        returns the request id; results via result(id) after wait(id)."""
        op = np.ascontiguousarray(op, np.uint32)
        a = np.ascontiguousarray(a, np.uint64)
        b = np.ascontiguousarray(b if b is not None else np.zeros(op.size), np.uint64)
        res = np.zeros(max(1, op.size), np.uint64)
        rid = ctypes.c_uint64()
        vp = ctypes.c_void_p
        _check(_lib.pwmhip_server_query(self.h, index, op.ctypes.data_as(vp), a.ctypes.data_as(vp),
                                        b.ctypes.data_as(vp), op.size, res.ctypes.data_as(vp),
                                        ctypes.byref(rid)), "server_query")
        self._keep[rid.value] = (op, a, b)
        self._results[rid.value] = res[:op.size]
        return rid.value

    def wait(self, rid, timeout_ms=-1):
        st = _lib.pwmhip_server_wait(self.h, rid, timeout_ms)
        if st != ERR_TIMEOUT:
            self._keep.pop(rid, None)
        _check(st, "server_wait")

    def result(self, rid):
        return self._results.pop(rid)

    def info(self, index):
        ws = _WS()
        _check(_lib.pwmhip_server_index_info(self.h, index, ctypes.byref(ws)), "server_index_info")
        return ws

    def levels(self, index):
        """Host copies of every level of a ready index."""
        ws = self.info(index)
        out = []
        for l in range(ws.levels):
            nw = (ws.level_bits[l] + 63) // 64
            a = np.zeros(max(1, nw), np.uint64)
            if nw:
                _check(_lib.pwmhip_server_download(self.h, index, l, a.ctypes.data_as(ctypes.c_void_p)),
                       "server_download")
            out.append(a[:nw])
        return out

    def zeros(self, index):
        ws = self.info(index)
        return np.array([ws.zeros[i] for i in range(ws.n_zeros)], np.uint64)

    def release(self, index):
        _check(_lib.pwmhip_server_release(self.h, index), "server_release")

    def stats(self):
        st = _ServerStats()
        _check(_lib.pwmhip_server_stats_get(self.h, ctypes.byref(st)), "server_stats")
        return {k: getattr(st, k) for k, _ in _ServerStats._fields_}
