"""Host-side mirror of the reference correction entry point, over the HIP C-ABI.

``correct_gpu(opts, storage)`` plays the role of the reference's
``care::cpu::correct_cpu(programOptions, minhasher, readStorage)``
(include/correct_cpu.hpp:17-21): given a read storage it builds the device-resident
minhash tables and returns the serialized correction records in read-id order.

The native library ``rabbitqcplus_amd/_lib/libcare_hip.so`` is required; there is no
CPU fallback. A missing library or device raises ``CareError``.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

from . import build as _build

LIB_PATH = os.path.join(_build.LIBDIR, "libcare_hip.so")
_LIB = None

CARE_OK = 0


class CareError(RuntimeError):
    pass


class CareOpts(ctypes.Structure):
    _fields_ = [("kmerlength", ctypes.c_int), ("num_hash_maps", ctypes.c_int),
                ("estimated_coverage", ctypes.c_float), ("use_quality_scores", ctypes.c_int),
                ("exclude_ambiguous", ctypes.c_int), ("paired", ctypes.c_int),
                ("min_overlap", ctypes.c_int), ("max_error_rate", ctypes.c_float),
                ("min_overlap_ratio", ctypes.c_float), ("estimated_errorrate", ctypes.c_float),
                ("m_coverage", ctypes.c_float), ("paired_filter_threshold", ctypes.c_float),
                ("batch_size", ctypes.c_int), ("keep_intermediates", ctypes.c_int),
                ("engine", ctypes.c_int), ("waves_per_cu", ctypes.c_int),
                ("correction_type", ctypes.c_int), ("align_variant", ctypes.c_int),
                ("use_graphs", ctypes.c_int), ("correct_candidates", ctypes.c_int),
                ("new_columns_to_correct", ctypes.c_int), ("msa_tile", ctypes.c_int),
                ("seeded_alignment", ctypes.c_int), ("sync_batches", ctypes.c_int),
                ("pipeline_capacity", ctypes.c_int)]


class CareTimings(ctypes.Structure):
    _fields_ = [("total_ms", ctypes.c_double), ("candidates_ms", ctypes.c_double),
                ("align_ms", ctypes.c_double), ("msa_ms", ctypes.c_double),
                ("output_ms", ctypes.c_double), ("align_kernel_ms", ctypes.c_double),
                ("cand_kernel_ms", ctypes.c_double), ("msa_kernel_ms", ctypes.c_double),
                ("fused_kernel_ms", ctypes.c_double),
                ("anchors", ctypes.c_uint64), ("candidates", ctypes.c_uint64),
                ("records", ctypes.c_uint64), ("record_bytes", ctypes.c_uint64),
                ("bucket_values", ctypes.c_uint64), ("batches", ctypes.c_uint64),
                ("graphs_built", ctypes.c_uint64), ("fastq_device_ms", ctypes.c_double),
                ("merge_device_ms", ctypes.c_double), ("graph_ms", ctypes.c_double),
                ("batches_redone", ctypes.c_uint64)]


EXPORTS = [
    "care_hip_default_opts", "care_hip_create", "care_hip_destroy", "care_hip_last_error",
    "care_hip_set_reads", "care_hip_encode_reads", "care_hip_build_tables", "care_hip_correct", "care_hip_copy_records",
    "care_hip_records_view",
    "care_hip_get_timings", "care_hip_inspect_signatures", "care_hip_inspect_batch",
    "care_hip_inspect_msa", "care_hip_set_classifier", "care_hip_extract_features", "care_hip_device_info",
    "care_hip_save_tables", "care_hip_load_tables", "care_hip_set_reads_fastq", "care_hip_read_layout",
    "care_hip_inspect_reads", "care_hip_write_fastq", "care_hip_save_reads", "care_hip_load_reads",
    "care_mlp_create", "care_mlp_destroy", "care_mlp_last_error", "care_mlp_set_params", "care_mlp_get_params",
    "care_mlp_reset_optimizer", "care_mlp_forward", "care_mlp_forward_static", "care_mlp_loss_grad",
    "care_mlp_set_data", "care_mlp_train",
]

MLP_PARAMS = 5313   # CARE_MLP_PARAMS: W1[64][16], b1[64], W2[64][64], b2[64], W3[64], b3


def load_library(path: str = LIB_PATH):
    """Loads the HIP library (no GPU call is made). Raises CareError if absent."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise CareError(f"HIP library not built: {path} (run __graft_entry__.build())")
    lib = ctypes.CDLL(path)
    vp, sz, u64, i32 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint64, ctypes.c_int
    lib.care_hip_default_opts.argtypes = [ctypes.POINTER(CareOpts)]
    lib.care_hip_default_opts.restype = None
    lib.care_hip_create.argtypes = [ctypes.POINTER(CareOpts), i32, ctypes.POINTER(vp)]
    lib.care_hip_destroy.argtypes = [vp]
    lib.care_hip_destroy.restype = None
    lib.care_hip_last_error.argtypes = [vp]
    lib.care_hip_last_error.restype = ctypes.c_char_p
    lib.care_hip_set_reads.argtypes = [vp, vp, sz, vp, vp, sz, u64, vp]
    lib.care_hip_set_reads_fastq.argtypes = [vp, vp, sz, vp, sz]
    lib.care_hip_encode_reads.argtypes = [vp, sz, vp, u64, vp, sz, vp, vp, sz, vp, sz, i32]
    lib.care_hip_read_layout.argtypes = [vp, ctypes.POINTER(u64), ctypes.POINTER(i32), ctypes.POINTER(sz)]
    lib.care_hip_inspect_reads.argtypes = [vp, u64, u64, vp, vp, vp, sz, vp]
    lib.care_hip_write_fastq.argtypes = [vp, vp, sz, i32, i32, vp, sz, ctypes.POINTER(sz)]
    lib.care_hip_save_reads.argtypes = [vp, ctypes.c_char_p]
    lib.care_hip_load_reads.argtypes = [vp, ctypes.c_char_p]
    lib.care_hip_build_tables.argtypes = [vp, ctypes.POINTER(i32)]
    lib.care_hip_correct.argtypes = [vp, u64, u64, ctypes.POINTER(sz), ctypes.POINTER(sz)]
    lib.care_hip_copy_records.argtypes = [vp, vp, sz]
    lib.care_hip_records_view.argtypes = [vp, ctypes.POINTER(vp), ctypes.POINTER(sz)]
    lib.care_hip_get_timings.argtypes = [vp, ctypes.POINTER(CareTimings)]
    lib.care_hip_inspect_signatures.argtypes = [vp, u64, u64, vp]
    lib.care_hip_inspect_batch.argtypes = [vp, ctypes.POINTER(u64), ctypes.POINTER(u64), ctypes.POINTER(u64),
                                           vp, vp, vp, vp, vp]
    lib.care_hip_inspect_msa.argtypes = [vp, u64, i32, ctypes.POINTER(i32), vp, vp, vp, vp, vp]
    lib.care_hip_device_info.argtypes = [i32, ctypes.c_char_p, sz, ctypes.POINTER(i32)]
    lib.care_hip_save_tables.argtypes = [vp, ctypes.c_char_p]
    lib.care_hip_load_tables.argtypes = [vp, ctypes.c_char_p, ctypes.POINTER(i32)]
    lib.care_hip_set_classifier.argtypes = [vp, vp, sz]
    lib.care_hip_extract_features.argtypes = [vp, u64, u64, vp, vp, sz, ctypes.POINTER(sz)]
    lib.care_mlp_create.argtypes = [i32, ctypes.POINTER(vp)]
    lib.care_mlp_destroy.argtypes = [vp]
    lib.care_mlp_destroy.restype = None
    lib.care_mlp_last_error.argtypes = [vp]
    lib.care_mlp_last_error.restype = ctypes.c_char_p
    lib.care_mlp_set_params.argtypes = [vp, vp, sz]
    lib.care_mlp_get_params.argtypes = [vp, vp, sz]
    lib.care_mlp_reset_optimizer.argtypes = [vp]
    lib.care_mlp_forward.argtypes = [vp, vp, sz, vp]
    lib.care_mlp_forward_static.argtypes = [vp, vp, sz, vp]
    lib.care_mlp_loss_grad.argtypes = [vp, vp, vp, sz, vp, vp]
    lib.care_mlp_set_data.argtypes = [vp, vp, vp, sz]
    lib.care_mlp_train.argtypes = [vp, i32, i32, ctypes.c_float, u64, vp]
    for name in EXPORTS:
        getattr(lib, name)
    _LIB = lib
    return lib


def device_info(device: int = 0) -> tuple[str, int]:
    """(device name and architecture, compute units)."""
    lib = load_library()
    buf = ctypes.create_string_buffer(256)
    cus = ctypes.c_int(0)
    rc = lib.care_hip_device_info(device, buf, 256, ctypes.byref(cus))
    if rc != CARE_OK:
        raise CareError(f"care_hip_device_info failed with code {rc}")
    return buf.value.decode(errors="replace"), cus.value


def default_opts(**kw) -> CareOpts:
    lib = load_library()
    o = CareOpts()
    lib.care_hip_default_opts(ctypes.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise KeyError(k)
        setattr(o, k, v)
    return o


@dataclass
class BatchInspection:
    first: int
    n: int
    offsets: np.ndarray
    candidate_ids: np.ndarray
    alignments: np.ndarray
    paired: np.ndarray
    anchor_info: np.ndarray


class CareEngine:
    """One device context: read storage + tables resident on the GPU."""

    def __init__(self, device: int = 0, **opts):
        self.lib = load_library()
        self.opts = default_opts(**opts)
        h = ctypes.c_void_p()
        rc = self.lib.care_hip_create(ctypes.byref(self.opts), device, ctypes.byref(h))
        if rc != CARE_OK:
            raise CareError(f"care_hip_create failed with code {rc}")
        self.h = h
        self.storage = None

    def _check(self, rc, what):
        if rc != CARE_OK:
            msg = self.lib.care_hip_last_error(self.h).decode(errors="replace")
            raise CareError(f"{what} failed ({rc}): {msg}")

    def close(self):
        if getattr(self, "h", None):
            self.lib.care_hip_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_reads(self, rs):
        self.storage = rs
        q = rs.quals if self.opts.use_quality_scores else None
        if self.opts.use_quality_scores and q is None:
            raise CareError("use_quality_scores set but the storage has no qualities")
        rc = self.lib.care_hip_set_reads(self.h, rs.enc.ctypes.data, rs.pitch_ints, rs.lengths.ctypes.data,
                                         None if q is None else q.ctypes.data, 0 if q is None else q.shape[1],
                                         rs.n, rs.ambiguous.ctypes.data)
        self._check(rc, "care_hip_set_reads")

    def set_reads_fastq(self, file1, file2=None) -> int:
        """Builds the read storage on the device from FASTQ files (paths or bytes);
        two files for paired input. Returns the number of reads."""
        def text(f):
            if isinstance(f, np.ndarray):
                return np.ascontiguousarray(f.reshape(-1).view(np.uint8))
            if isinstance(f, (bytes, bytearray, memoryview)):
                return np.frombuffer(f, np.uint8)
            return np.fromfile(f, np.uint8)
        t1 = text(file1)
        t2 = None if file2 is None else text(file2)
        self.storage = None
        self._fastq_bytes = [t1.size] + ([] if t2 is None else [t2.size])
        rc = self.lib.care_hip_set_reads_fastq(self.h, t1.ctypes.data, t1.size,
                                               None if t2 is None else t2.ctypes.data, 0 if t2 is None else t2.size)
        self._check(rc, "care_hip_set_reads_fastq")
        return self.read_layout()[0]

    def write_fastq(self, records: bytes, file: int = 0, quality_labels: bool = False, as_array: bool = False):
        """Corrected FASTQ text of input file `file` from serialized records (output merge
        on the device; needs reads set by set_reads_fastq): bytes, or with as_array a uint8
        array holding the text without a further copy (to write to a file, say)."""
        rec = np.frombuffer(records, np.uint8) if len(records) else np.zeros(1, np.uint8)
        size = ctypes.c_size_t(0)
        # every read's output is at most its input text, a label and its records' bytes
        fq = getattr(self, "_fastq_bytes", None)
        if fq is not None and 0 <= file < len(fq):
            bound = fq[file] + 9 * (self.num_reads // len(fq)) + len(records) + 64
        else:
            self._check(self.lib.care_hip_write_fastq(self.h, rec.ctypes.data, len(records), file,
                                                      int(quality_labels), None, 0, ctypes.byref(size)),
                        "care_hip_write_fastq")
            bound = size.value
        out = np.empty(max(bound, 1), np.uint8)
        self._check(self.lib.care_hip_write_fastq(self.h, rec.ctypes.data, len(records), file, int(quality_labels),
                                                  out.ctypes.data, out.size, ctypes.byref(size)), "care_hip_write_fastq")
        return out[:size.value] if as_array else out[:size.value].tobytes()

    def save_reads(self, path: str) -> None:
        """Writes the resident read storage as the reference's preprocessed-reads file."""
        self._check(self.lib.care_hip_save_reads(self.h, os.fsencode(path)), "care_hip_save_reads")

    def load_reads(self, path: str) -> int:
        """Loads the read storage from a preprocessed-reads file; returns the number of reads."""
        self.storage = None
        self._check(self.lib.care_hip_load_reads(self.h, os.fsencode(path)), "care_hip_load_reads")
        return self.read_layout()[0]

    def read_layout(self) -> tuple[int, int, int]:
        """(number of reads, longest read, pitch in u32 words) of the resident storage."""
        n, ml, pitch = ctypes.c_uint64(0), ctypes.c_int(0), ctypes.c_size_t(0)
        self._check(self.lib.care_hip_read_layout(self.h, ctypes.byref(n), ctypes.byref(ml), ctypes.byref(pitch)),
                    "care_hip_read_layout")
        return n.value, ml.value, pitch.value

    @property
    def num_reads(self) -> int:
        return self.read_layout()[0]

    def inspect_reads(self, first: int = 0, n: int | None = None) -> dict:
        """Reads [first, first+n) of the resident storage: enc [n, pitch] u32, lengths,
        quals [n, longest] u8 (None without quality scores), ambiguous."""
        total, ml, pitch = self.read_layout()
        n = total - first if n is None else n
        enc = np.zeros((max(n, 1), pitch), np.uint32)
        lens = np.zeros(max(n, 1), np.int32)
        amb = np.zeros(max(n, 1), np.uint8)
        quals = np.zeros((max(n, 1), ml), np.uint8) if self.opts.use_quality_scores else None
        self._check(self.lib.care_hip_inspect_reads(self.h, first, n, enc.ctypes.data, lens.ctypes.data,
                                                    None if quals is None else quals.ctypes.data, ml,
                                                    amb.ctypes.data), "care_hip_inspect_reads")
        return dict(enc=enc[:n], lengths=lens[:n], quals=None if quals is None else quals[:n], ambiguous=amb[:n],
                    max_length=ml)

    def build_tables(self) -> int:
        maps = ctypes.c_int(0)
        self._check(self.lib.care_hip_build_tables(self.h, ctypes.byref(maps)), "care_hip_build_tables")
        return maps.value

    def save_tables(self, path: str) -> None:
        """Writes the tables in the reference's minhasher file format."""
        self._check(self.lib.care_hip_save_tables(self.h, os.fsencode(path)), "care_hip_save_tables")

    def load_tables(self, path: str) -> int:
        """Loads tables from a reference minhasher file (instead of build_tables)."""
        maps = ctypes.c_int(0)
        self._check(self.lib.care_hip_load_tables(self.h, os.fsencode(path), ctypes.byref(maps)), "care_hip_load_tables")
        return maps.value

    def correct(self, first: int = 0, n: int | None = None) -> tuple[int, bytes]:
        n = self.num_reads - first if n is None else n
        nrec, nbytes = ctypes.c_size_t(0), ctypes.c_size_t(0)
        self._check(self.lib.care_hip_correct(self.h, first, n, ctypes.byref(nrec), ctypes.byref(nbytes)),
                    "care_hip_correct")
        # one copy, from the context's pinned record buffer into the returned bytes
        ptr, size = ctypes.c_void_p(), ctypes.c_size_t(0)
        self._check(self.lib.care_hip_records_view(self.h, ctypes.byref(ptr), ctypes.byref(size)), "care_hip_records_view")
        return nrec.value, ctypes.string_at(ptr.value, size.value) if size.value else b""

    def set_classifier(self, params) -> None:
        """MLP parameters (MLP_PARAMS float32 values) for correction_type=1; None unloads."""
        if params is None:
            self._check(self.lib.care_hip_set_classifier(self.h, None, 0), "care_hip_set_classifier")
            return
        p = np.ascontiguousarray(params, np.float32)
        self._check(self.lib.care_hip_set_classifier(self.h, p.ctypes.data, p.size), "care_hip_set_classifier")

    def extract_features(self, first: int = 0, n: int | None = None, cap: int | None = None):
        """Decision columns of anchors [first, first+n): (features [m,16] f32, meta [m,4] u32),
        ordered by (read id, position)."""
        n = self.num_reads - first if n is None else n
        total = ctypes.c_size_t(0)
        if cap is None:
            self._check(self.lib.care_hip_extract_features(self.h, first, n, None, None, 0, ctypes.byref(total)),
                        "care_hip_extract_features")
            cap = total.value
        feats = np.zeros((max(cap, 1), 16), np.float32)
        meta = np.zeros((max(cap, 1), 4), np.uint32)
        self._check(self.lib.care_hip_extract_features(self.h, first, n, feats.ctypes.data, meta.ctypes.data, cap,
                                                       ctypes.byref(total)), "care_hip_extract_features")
        m = min(total.value, cap)
        return feats[:m], meta[:m]

    def timings(self) -> dict:
        t = CareTimings()
        self._check(self.lib.care_hip_get_timings(self.h, ctypes.byref(t)), "care_hip_get_timings")
        return {k: getattr(t, k) for k, _ in CareTimings._fields_}

    def signatures(self, first: int = 0, n: int | None = None) -> np.ndarray:
        n = self.num_reads - first if n is None else n
        out = np.zeros((n, self.opts.num_hash_maps), np.uint64)
        self._check(self.lib.care_hip_inspect_signatures(self.h, first, n, out.ctypes.data), "inspect_signatures")
        return out

    def inspect_batch(self) -> BatchInspection:
        f, b, t = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        self._check(self.lib.care_hip_inspect_batch(self.h, ctypes.byref(f), ctypes.byref(b), ctypes.byref(t),
                                                    None, None, None, None, None), "inspect_batch")
        B, T = b.value, t.value
        offs = np.zeros(B + 1, np.uint32)
        ids = np.zeros(max(T, 1), np.uint32)
        al = np.zeros((max(T, 1), 11), np.int32)
        pr = np.zeros(max(T, 1), np.uint8)
        info = np.zeros((B, 12), np.int32)
        self._check(self.lib.care_hip_inspect_batch(self.h, None, None, None, offs.ctypes.data, ids.ctypes.data,
                                                    al.ctypes.data, pr.ctypes.data, info.ctypes.data), "inspect_batch")
        return BatchInspection(f.value, B, offs, ids[:T], al[:T], pr[:T], info)

    def inspect_msa(self, index: int, cap_columns: int = 4096) -> dict:
        ncol = ctypes.c_int(0)
        counts = np.zeros((cap_columns, 4), np.int32)
        weights = np.zeros((cap_columns, 4), np.float32)
        cov = np.zeros(cap_columns, np.int32)
        cons = np.zeros(cap_columns, np.uint8)
        sup = np.zeros(cap_columns, np.float32)
        self._check(self.lib.care_hip_inspect_msa(self.h, index, cap_columns, ctypes.byref(ncol), counts.ctypes.data,
                                                  weights.ctypes.data, cov.ctypes.data, cons.ctypes.data,
                                                  sup.ctypes.data), "inspect_msa")
        c = ncol.value
        return dict(num_columns=c, counts=counts[:c], weights=weights[:c], coverage=cov[:c],
                    consensus=cons[:c], support=sup[:c])


def correct_gpu(storage, device: int = 0, classifier=None, **opts) -> tuple[int, bytes]:
    """Reads -> tables -> records, the whole correction stage on one GPU. With a
    classifier (MLP parameters) the learned seam is used (correction_type=1)."""
    opts.setdefault("paired", int(storage.paired))
    opts.setdefault("use_quality_scores", int(storage.quals is not None))
    if classifier is not None:
        opts.setdefault("correction_type", 1)
    eng = CareEngine(device, **opts)
    try:
        eng.set_reads(storage)
        eng.build_tables()
        if classifier is not None:
            eng.set_classifier(classifier)
        return eng.correct()
    finally:
        eng.close()
