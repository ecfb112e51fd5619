"""Python handles on the two CPU checkers (test infrastructure only).

``CareCpu``  - our C++ restatement (oracle/carecpu.cpp), always available once built.
``CareRef``  - the reference engine compiled from its own sources (oracle/_ref),
               available where it was built; it never reads the reference tree at
               run time.

Only tests, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg use
these, and only as the checker.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
CPU_LIB = os.path.join(HERE, "_build", "libcarecpu.so")
REF_LIB = os.path.join(HERE, "_ref", "libcareref.so")

_cpu = None
_ref = None

DEFAULTS = dict(kmerlength=20, num_hash_maps=48, estimated_coverage=30.0, use_quality_scores=0,
                exclude_ambiguous=0, paired=0, min_overlap=30, max_error_rate=0.2,
                min_overlap_ratio=0.3, estimated_errorrate=0.06, m_coverage=0.6,
                paired_filter_threshold=0.06)


class CpuOpts(ctypes.Structure):
    _fields_ = [("kmerlength", ctypes.c_int), ("num_hash_maps", ctypes.c_int),
                ("estimated_coverage", ctypes.c_float), ("use_quality_scores", ctypes.c_int),
                ("exclude_ambiguous", ctypes.c_int), ("paired", ctypes.c_int),
                ("min_overlap", ctypes.c_int), ("max_error_rate", ctypes.c_float),
                ("min_overlap_ratio", ctypes.c_float), ("estimated_errorrate", ctypes.c_float),
                ("m_coverage", ctypes.c_float), ("paired_filter_threshold", ctypes.c_float)]


class RefOpts(ctypes.Structure):
    _fields_ = [("kmerlength", ctypes.c_int), ("num_hash_maps", ctypes.c_int),
                ("estimated_coverage", ctypes.c_float), ("threads", ctypes.c_int),
                ("use_quality_scores", ctypes.c_int), ("exclude_ambiguous", ctypes.c_int),
                ("paired", ctypes.c_int), ("min_overlap", ctypes.c_int),
                ("max_error_rate", ctypes.c_float), ("min_overlap_ratio", ctypes.c_float),
                ("estimated_errorrate", ctypes.c_float), ("m_coverage", ctypes.c_float),
                ("paired_filter_threshold", ctypes.c_float), ("fixed_number_of_reads", ctypes.c_longlong)]


class Trace(ctypes.Structure):
    _fields_ = [("cap_candidates", ctypes.c_int), ("cap_columns", ctypes.c_int),
                ("num_candidates", ctypes.c_int), ("candidate_ids", ctypes.c_void_p),
                ("fwd_alignments", ctypes.c_void_p), ("rc_alignments", ctypes.c_void_p),
                ("orientation", ctypes.c_void_p), ("is_paired", ctypes.c_void_p),
                ("num_filtered", ctypes.c_int), ("num_final", ctypes.c_int),
                ("final_ids", ctypes.c_void_p), ("final_shifts", ctypes.c_void_p),
                ("final_weights", ctypes.c_void_p), ("refinement_rounds", ctypes.c_int),
                ("num_columns", ctypes.c_int), ("anchor_column_begin", ctypes.c_int),
                ("counts", ctypes.c_void_p), ("weights", ctypes.c_void_p), ("coverage", ctypes.c_void_p),
                ("consensus", ctypes.c_void_p), ("support", ctypes.c_void_p),
                ("orig_weights", ctypes.c_void_p), ("orig_coverage", ctypes.c_void_p),
                ("avg_support", ctypes.c_float), ("min_support", ctypes.c_float),
                ("min_coverage", ctypes.c_int), ("max_coverage", ctypes.c_int),
                ("is_hq", ctypes.c_int), ("is_corrected", ctypes.c_int), ("corrected", ctypes.c_void_p)]


def build():
    subprocess.run(["make", "-s", "-C", HERE, "carecpu"], check=True)


def _p(a):
    return None if a is None else a.ctypes.data


def cpu_lib():
    global _cpu
    if _cpu is None:
        if not os.path.exists(CPU_LIB):
            build()
        lib = ctypes.CDLL(CPU_LIB)
        vp, sz, u64, i32 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint64, ctypes.c_int
        lib.carecpu_create.argtypes = [ctypes.POINTER(CpuOpts), u64, vp, sz, vp, vp, sz, vp, i32]
        lib.carecpu_create.restype = vp
        lib.carecpu_destroy.argtypes = [vp]
        lib.carecpu_results_per_map_threshold.argtypes = [ctypes.c_float]
        lib.carecpu_signature.argtypes = [vp, i32, i32, i32, vp]
        lib.carecpu_signatures.argtypes = [vp, u64, u64, vp]
        lib.carecpu_bucket.argtypes = [vp, i32, u64, ctypes.POINTER(ctypes.c_void_p)]
        lib.carecpu_candidates.argtypes = [vp, ctypes.c_uint32, vp, i32]
        lib.carecpu_align.argtypes = [vp, i32, vp, i32, i32, ctypes.c_float, ctypes.c_float, vp]
        lib.carecpu_orientation.argtypes = [vp, vp, i32, i32, ctypes.c_float, i32, ctypes.c_float]
        lib.carecpu_trace_anchor.argtypes = [vp, ctypes.c_uint32, ctypes.POINTER(Trace)]
        lib.carecpu_correct.argtypes = [vp, u64, u64, i32]
        lib.carecpu_correct.restype = sz
        lib.carecpu_records_bytes.argtypes = [vp]
        lib.carecpu_records_bytes.restype = sz
        lib.carecpu_records_data.argtypes = [vp]
        lib.carecpu_records_data.restype = vp
        lib.carecpu_last_seconds.argtypes = [vp]
        lib.carecpu_flag_pairs.argtypes = [vp, i32, vp, i32, vp, vp]
        lib.carecpu_set_classifier.argtypes = [vp, vp, i32]
        lib.carecpu_set_candidate_correction.argtypes = [vp, i32, i32]
        lib.carecpu_set_candidate_correction.restype = None
        lib.carecpu_set_seeded_alignment.argtypes = [vp, i32]
        lib.carecpu_set_seeded_alignment.restype = None
        lib.carecpu_seeds.argtypes = [vp, u64, u64, vp]
        lib.carecpu_seeds.restype = None
        lib.carecpu_seeded_alignment.argtypes = [vp, ctypes.c_uint32, ctypes.c_uint32, vp]
        lib.carecpu_seeded_alignment.restype = None
        lib.carecpu_set_classifier.restype = i32
        lib.carecpu_mlp_num_params.restype = i32
        lib.carecpu_features.argtypes = [vp, u64, u64, i32, vp, vp, sz]
        lib.carecpu_features.restype = sz
        lib.carecpu_mlp_logits_bf16.argtypes = [vp, vp, u64, vp]
        lib.carecpu_mlp_loss_grad_f64.argtypes = [vp, vp, vp, u64, vp, vp]
        lib.carecpu_mlp_loss_grad_f64.restype = ctypes.c_double
        lib.carecpu_last_seconds.restype = ctypes.c_double
        _cpu = lib
    return _cpu


def ref_available() -> bool:
    return os.path.exists(REF_LIB)


def ref_lib():
    global _ref
    if _ref is None:
        lib = ctypes.CDLL(REF_LIB)
        vp, sz, i32 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int
        lib.ref_create.argtypes = [ctypes.POINTER(RefOpts), sz, vp, sz, vp, vp, sz, vp]
        lib.ref_create.restype = vp
        lib.ref_create_loaded.argtypes = [ctypes.POINTER(RefOpts), sz, vp, sz, vp, vp, sz, vp, ctypes.c_char_p,
                                          ctypes.POINTER(ctypes.c_double)]
        lib.ref_create_loaded.restype = vp
        lib.ref_destroy.argtypes = [vp]
        lib.ref_num_maps.argtypes = [vp]
        lib.ref_correct.argtypes = [vp, i32, ctypes.c_longlong]
        lib.ref_correct.restype = sz
        lib.ref_correct_timed.argtypes = [vp, i32, ctypes.c_longlong, ctypes.POINTER(ctypes.c_longlong)]
        lib.ref_correct_timed.restype = ctypes.c_double
        lib.ref_last_seconds.argtypes = [vp]
        lib.ref_last_seconds.restype = ctypes.c_double
        lib.ref_records_bytes.argtypes = [vp]
        lib.ref_records_bytes.restype = sz
        lib.ref_records_data.argtypes = [vp]
        lib.ref_records_data.restype = vp
        lib.ref_hash.argtypes = [vp, i32, i32, i32, vp]
        lib.ref_candidates.argtypes = [vp, ctypes.c_uint, vp, i32]
        lib.ref_align.argtypes = [vp, i32, vp, i32, i32, ctypes.c_float, ctypes.c_float, vp]
        lib.ref_flag_pairs.argtypes = [vp, i32, vp, i32, vp, vp]
        lib.ref_set_anchor_forest.argtypes = [vp, ctypes.c_char_p, ctypes.c_float]
        lib.ref_set_anchor_print.argtypes = [vp, ctypes.c_char_p]
        lib.ref_set_candidate_correction.argtypes = [vp, i32, i32]
        lib.ref_set_candidate_correction.restype = None
        lib.ref_set_record_order.argtypes = [vp, i32]
        lib.ref_set_record_order.restype = None
        lib.ref_save_tables.argtypes = [vp, ctypes.c_char_p]
        lib.ref_parse_fastq.argtypes = [ctypes.c_char_p, ctypes.c_char_p, i32]
        lib.ref_parse_fastq.restype = vp
        lib.ref_reads_count.argtypes = [vp]
        lib.ref_reads_count.restype = sz
        lib.ref_reads_max_length.argtypes = [vp]
        lib.ref_reads_max_length.restype = i32
        lib.ref_reads_pitch.argtypes = [vp]
        lib.ref_reads_pitch.restype = sz
        lib.ref_reads_copy.argtypes = [vp, vp, vp, vp, vp]
        lib.ref_reads_free.argtypes = [vp]
        lib.ref_reads_save.argtypes = [vp, ctypes.c_char_p]
        lib.ref_reads_load.argtypes = [ctypes.c_char_p, i32, i32]
        lib.ref_reads_load.restype = vp
        lib.ref_output_fastq.argtypes = [ctypes.c_char_p, ctypes.c_char_p, vp, sz, i32, ctypes.c_char_p,
                                         ctypes.c_char_p, ctypes.c_char_p]
        lib.ref_load_tables.argtypes = [vp, ctypes.c_char_p]
        lib.ref_orientation.argtypes = [i32, i32, i32, i32, i32, i32, i32, i32, ctypes.c_float, i32, ctypes.c_float]
        lib.ref_msa.argtypes = [i32, i32, i32, vp, vp, vp, i32, vp, i32, vp, vp, vp] + [vp] * 8
        _ref = lib
    return _ref


def _opts(cls, **kw):
    o = cls()
    d = dict(DEFAULTS)
    d.update(kw)
    for name, _ in cls._fields_:
        if name in d:
            setattr(o, name, d[name])
    return o


class CareCpu:
    """Our CPU restatement over a ReadStorage."""

    def __init__(self, rs, threads: int = 0, **opts):
        self.lib = cpu_lib()
        self.rs = rs
        opts.setdefault("paired", int(rs.paired))
        opts.setdefault("use_quality_scores", int(rs.quals is not None))
        self.opts = _opts(CpuOpts, **opts)
        q = rs.quals
        self.h = self.lib.carecpu_create(ctypes.byref(self.opts), rs.n, rs.enc.ctypes.data, rs.pitch_ints,
                                         rs.lengths.ctypes.data, _p(q), 0 if q is None else q.shape[1],
                                         rs.ambiguous.ctypes.data, threads)
        if not self.h:
            raise RuntimeError("carecpu_create failed")

    def close(self):
        if self.h:
            self.lib.carecpu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def signatures(self, first=0, n=None):
        n = self.rs.n - first if n is None else n
        out = np.zeros((n, self.opts.num_hash_maps), np.uint64)
        self.lib.carecpu_signatures(self.h, first, n, out.ctypes.data)
        return out

    def candidates(self, read_id):
        cap = 48 * 1024
        buf = np.zeros(cap, np.uint32)
        n = self.lib.carecpu_candidates(self.h, read_id, buf.ctypes.data, cap)
        return buf[:n].copy()

    def trace(self, read_id, cap_candidates=8192, cap_columns=4096):
        t = Trace()
        bufs = dict(candidate_ids=np.zeros(cap_candidates, np.uint32),
                    fwd_alignments=np.zeros((cap_candidates, 5), np.int32),
                    rc_alignments=np.zeros((cap_candidates, 5), np.int32),
                    orientation=np.zeros(cap_candidates, np.int8),
                    is_paired=np.zeros(cap_candidates, np.uint8),
                    final_ids=np.zeros(cap_candidates, np.uint32),
                    final_shifts=np.zeros(cap_candidates, np.int32),
                    final_weights=np.zeros(cap_candidates, np.float32),
                    counts=np.zeros((cap_columns, 4), np.int32),
                    weights=np.zeros((cap_columns, 4), np.float32),
                    coverage=np.zeros(cap_columns, np.int32),
                    consensus=np.zeros(cap_columns, np.uint8),
                    support=np.zeros(cap_columns, np.float32),
                    orig_weights=np.zeros(cap_columns, np.float32),
                    orig_coverage=np.zeros(cap_columns, np.int32),
                    corrected=np.zeros(int(self.rs.lengths[read_id]), np.uint8))
        t.cap_candidates = cap_candidates
        t.cap_columns = cap_columns
        for k, v in bufs.items():
            setattr(t, k, v.ctypes.data)
        rc = self.lib.carecpu_trace_anchor(self.h, read_id, ctypes.byref(t))
        if rc != 0:
            raise RuntimeError(f"trace failed {rc}")
        nc, nf, ncol = t.num_candidates, t.num_final, t.num_columns
        res = {k: getattr(t, k) for k in ("num_candidates", "num_filtered", "num_final", "refinement_rounds",
                                          "num_columns", "anchor_column_begin", "avg_support", "min_support",
                                          "min_coverage", "max_coverage", "is_hq", "is_corrected")}
        for k in ("candidate_ids", "fwd_alignments", "rc_alignments", "orientation", "is_paired"):
            res[k] = bufs[k][:nc].copy()
        for k in ("final_ids", "final_shifts", "final_weights"):
            res[k] = bufs[k][:nf].copy()
        for k in ("counts", "weights", "coverage", "consensus", "support", "orig_weights", "orig_coverage"):
            res[k] = bufs[k][:ncol].copy()
        res["corrected"] = bufs["corrected"].tobytes() if t.is_corrected else None
        return res

    def correct(self, first=0, n=None, threads=0):
        n = self.rs.n - first if n is None else n
        cnt = self.lib.carecpu_correct(self.h, first, n, threads)
        nbytes = self.lib.carecpu_records_bytes(self.h)
        data = ctypes.string_at(self.lib.carecpu_records_data(self.h), nbytes) if nbytes else b""
        return cnt, data

    def last_seconds(self):
        return self.lib.carecpu_last_seconds(self.h)

    def set_candidate_correction(self, on=True, new_columns=15):
        self.lib.carecpu_set_candidate_correction(self.h, int(on), int(new_columns))

    def set_seeded_alignment(self, on=True):
        """Seeded alignment (the fast learned form's alignment) instead of the shift scan."""
        self.lib.carecpu_set_seeded_alignment(self.h, int(on))

    def seeds(self, first=0, n=None):
        n = self.rs.n - first if n is None else n
        out = np.zeros((n, self.opts.num_hash_maps), np.uint16)
        self.lib.carecpu_seeds(self.h, first, n, out.ctypes.data)
        return out

    def seeded_alignment(self, anchor, candidate):
        """(map or -1, flag 1/2/3, shift, overlap, mismatches) of the seeded alignment."""
        out = np.zeros(5, np.int32)
        self.lib.carecpu_seeded_alignment(self.h, int(anchor), int(candidate), out.ctypes.data)
        return tuple(int(x) for x in out)

    def set_classifier(self, params):
        """MLP parameters (float32, MLP_PARAMS values) for the learned seam; None = classic."""
        if params is None:
            self.lib.carecpu_set_classifier(self.h, None, 0)
            return
        self._params = np.ascontiguousarray(params, np.float32)
        if self.lib.carecpu_set_classifier(self.h, self._params.ctypes.data, self._params.size) != 0:
            raise ValueError("wrong number of classifier parameters")

    def features(self, first=0, n=None, threads=0):
        """Decision columns of anchors [first, first+n): (features [m,16] f32, meta [m,4] u32)."""
        n = self.rs.n - first if n is None else n
        total = self.lib.carecpu_features(self.h, first, n, threads, None, None, 0)
        feats = np.zeros((max(total, 1), 16), np.float32)
        meta = np.zeros((max(total, 1), 4), np.uint32)
        self.lib.carecpu_features(self.h, first, n, threads, feats.ctypes.data, meta.ctypes.data, total)
        return feats[:total], meta[:total]


DROPIN_LIB = os.path.join(HERE, "_ref", "libcaredropin.so")
_dropin = None


def dropin_available() -> bool:
    return os.path.exists(DROPIN_LIB)


def dropin_correct(rs, device: int = 0, **opts):
    """The reference-side drop-in care::mi355x::correct (rabbitqcplus_amd/dropin/, compiled
    against the reference headers, linked to libcare_hip.so) over the harness's
    CpuReadStorage: (record count, the bytes of its SerializedObjectStorage in order)."""
    global _dropin
    if _dropin is None:
        lib = ctypes.CDLL(DROPIN_LIB)
        vp, sz = ctypes.c_void_p, ctypes.c_size_t
        lib.ref_storage.argtypes = [ctypes.POINTER(RefOpts), sz, vp, sz, vp, vp, sz, vp]
        lib.ref_storage.restype = vp
        lib.ref_destroy.argtypes = [vp]
        lib.dropin_correct.argtypes = [vp, ctypes.c_int, ctypes.c_char_p]
        lib.dropin_correct.restype = ctypes.c_longlong
        lib.ref_records_bytes.argtypes = [vp]
        lib.ref_records_bytes.restype = sz
        lib.ref_records_data.argtypes = [vp]
        lib.ref_records_data.restype = vp
        _dropin = lib
    lib = _dropin
    opts.setdefault("paired", int(rs.paired))
    opts.setdefault("use_quality_scores", int(rs.quals is not None))
    opts.setdefault("threads", 8)
    opts.setdefault("fixed_number_of_reads", 0)
    o = _opts(RefOpts, **opts)
    q = rs.quals
    h = lib.ref_storage(ctypes.byref(o), rs.n, rs.enc.ctypes.data, rs.pitch_ints, rs.lengths.ctypes.data, _p(q),
                        0 if q is None else q.shape[1], rs.ambiguous.ctypes.data)
    if not h:
        raise RuntimeError("ref_storage failed")
    try:
        err = ctypes.create_string_buffer(256)
        n = lib.dropin_correct(h, device, err)
        if n < 0:
            raise RuntimeError(err.value.decode(errors="replace"))
        nb = lib.ref_records_bytes(h)
        return int(n), (ctypes.string_at(lib.ref_records_data(h), nb) if nb else b"")
    finally:
        lib.ref_destroy(h)


class CareRef:
    """The reference engine (compiled from source) over the same ReadStorage."""

    def __init__(self, rs, threads: int = 8, tables: str | None = None, **opts):
        """tables: a minhasher file in the reference's format, read by the reference's own
        loader instead of building the tables (self.load_seconds = its wall time)."""
        self.lib = ref_lib()
        self.rs = rs
        opts.setdefault("paired", int(rs.paired))
        opts.setdefault("use_quality_scores", int(rs.quals is not None))
        opts.setdefault("threads", threads)
        opts.setdefault("fixed_number_of_reads", 0)
        self.opts = _opts(RefOpts, **opts)
        q = rs.quals
        args = (ctypes.byref(self.opts), rs.n, rs.enc.ctypes.data, rs.pitch_ints, rs.lengths.ctypes.data, _p(q),
                0 if q is None else q.shape[1], rs.ambiguous.ctypes.data)
        self.load_seconds = None
        if tables is None:
            self.h = self.lib.ref_create(*args)
        else:
            secs = ctypes.c_double(0.0)
            self.h = self.lib.ref_create_loaded(*args, os.fsencode(tables), ctypes.byref(secs))
            self.load_seconds = secs.value
        if not self.h:
            raise RuntimeError("ref_create failed")

    def close(self):
        if self.h:
            self.lib.ref_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_candidate_correction(self, on=True, new_columns=15):
        """Candidate correction; run correct(threads=1) for the deterministic order."""
        self.lib.ref_set_candidate_correction(self.h, int(on), int(new_columns))

    def set_record_order(self, sorted_by_read_id=True):
        """Records of correct() sorted by read id (default) or in insertion order."""
        self.lib.ref_set_record_order(self.h, int(sorted_by_read_id))

    def correct(self, threads=8, num=0):
        cnt = self.lib.ref_correct(self.h, threads, num)
        nbytes = self.lib.ref_records_bytes(self.h)
        data = ctypes.string_at(self.lib.ref_records_data(self.h), nbytes) if nbytes else b""
        return cnt, data

    def correct_timed(self, threads, num_anchors):
        recs = ctypes.c_longlong(0)
        secs = self.lib.ref_correct_timed(self.h, threads, num_anchors, ctypes.byref(recs))
        return secs, recs.value

    def candidates(self, read_id):
        cap = 48 * 1024
        buf = np.zeros(cap, np.uint32)
        n = self.lib.ref_candidates(self.h, read_id, buf.ctypes.data, cap)
        return buf[:n].copy()

    def print_anchor_features(self, path):
        """Runs the reference's feature print mode over all anchors (one thread) and returns
        (read ids, positions, consensus chars, features [m,13]) parsed from `path`."""
        self.lib.ref_set_anchor_print(self.h, os.fsencode(path))
        self.lib.ref_correct(self.h, 1, 0)
        self.lib.ref_set_anchor_forest(self.h, None, 0.5)
        rows = []
        with open(path) as fh:
            fh.readline()                        # extractor descriptor
            for line in fh:
                parts = line.split()
                if parts:
                    rows.append(parts)
        rid = np.array([int(r[0]) for r in rows], np.int64)
        pos = np.array([int(r[1]) for r in rows], np.int64)
        cons = np.array([r[2] for r in rows])
        feats = np.array([[float(v) for v in r[3:]] for r in rows], np.float64).reshape(len(rows), -1)
        return rid, pos, cons, feats

    def save_tables(self, path):
        if self.lib.ref_save_tables(self.h, os.fsencode(path)) != 0:
            raise RuntimeError("ref_save_tables failed")

    def load_tables(self, path):
        n = self.lib.ref_load_tables(self.h, os.fsencode(path))
        if n < 0:
            raise RuntimeError("ref_load_tables failed")
        return n

    def set_anchor_forest(self, path, threshold=0.5):
        """Forest-classifier anchor correction with the forest file at `path`; None = classic."""
        self.lib.ref_set_anchor_forest(self.h, None if path is None else os.fsencode(path), threshold)


def write_constant_forest(path, prob):
    """A forest file of one single-leaf tree that answers `prob` for every column, in the
    reference's format (include/forest.hpp:141-162 and include/deserialize.hpp:31-36):
    u64 descriptor length, descriptor, u32 tree count, per tree u32 node count and the
    nodes in preorder (u8 feature, f64 split, u8 leaf flags, f32 leaf probabilities)."""
    import struct
    desc = b"13 extract_anchor_v3"
    blob = struct.pack("<Q", len(desc)) + desc + struct.pack("<I", 1) + struct.pack("<I", 1)
    blob += struct.pack("<B", 0) + struct.pack("<d", 0.0) + struct.pack("<B", 3) + struct.pack("<ff", prob, prob)
    with open(path, "wb") as fh:
        fh.write(blob)


def ref_output_fastq(file1, file2, records: bytes, labels=False, tmpdir=None):
    """The reference's output construction (sort + merge into QC's queues): ([text per
    queue], number of records). Single-end runs have one queue."""
    import tempfile
    lib = ref_lib()
    with tempfile.TemporaryDirectory(dir=tmpdir) as d:
        q1, q2 = os.path.join(d, "q1.out"), os.path.join(d, "q2.out")
        buf = ctypes.create_string_buffer(records, max(1, len(records)))
        rc = lib.ref_output_fastq(os.fsencode(file1), None if file2 is None else os.fsencode(file2), buf,
                                  len(records), int(labels), os.fsencode(d), os.fsencode(q1), os.fsencode(q2))
        if rc < 0:
            raise RuntimeError("ref_output_fastq failed")
        outs = [open(q1, "rb").read()]
        if file2 is not None:
            outs.append(open(q2, "rb").read())
    return outs, rc


def sort_records(data: bytes) -> bytes:
    """Serialized records ordered by read id, ties in stream order."""
    spans = []
    p = 0
    while p < len(data):
        flags = int.from_bytes(data[p + 4:p + 8], "little")
        nb = 8 + (flags & ((1 << 29) - 1))
        spans.append((int.from_bytes(data[p:p + 4], "little"), p, nb))
        p += nb
    spans.sort(key=lambda x: x[0])
    return b"".join(data[q:q + nb] for _, q, nb in spans)


def ref_parse_fastq(file1, file2=None, use_qualities=True, save_to=None):
    """The reference's FASTQ -> read storage construction: dict of enc [n, pitch] u32,
    lengths [n] i32, quals [n, maxLen] u8 (or None), ambiguous [n] u8, max_length.
    save_to: also write the reference's preprocessed-reads file there."""
    lib = ref_lib()
    h = lib.ref_parse_fastq(os.fsencode(file1), None if file2 is None else os.fsencode(file2), int(use_qualities))
    if save_to is not None and lib.ref_reads_save(h, os.fsencode(save_to)) != 0:
        lib.ref_reads_free(h)
        raise RuntimeError("ref_reads_save failed")
    return _reads_dict(lib, h, use_qualities)


def ref_load_reads(path, paired=False, use_qualities=True):
    """Read storage from a preprocessed-reads file, by the reference's loader."""
    lib = ref_lib()
    h = lib.ref_reads_load(os.fsencode(path), int(paired), int(use_qualities))
    if not h:
        raise RuntimeError(f"the reference could not load {path}")
    return _reads_dict(lib, h, use_qualities)


def _reads_dict(lib, h, use_qualities):
    try:
        n = lib.ref_reads_count(h)
        ml = lib.ref_reads_max_length(h)
        pitch = lib.ref_reads_pitch(h)
        enc = np.zeros((n, pitch), np.uint32)
        lens = np.zeros(n, np.int32)
        quals = np.zeros((n, ml), np.uint8) if use_qualities else None
        amb = np.zeros(n, np.uint8)
        lib.ref_reads_copy(h, enc.ctypes.data, lens.ctypes.data, _p(quals), amb.ctypes.data)
    finally:
        lib.ref_reads_free(h)
    return dict(enc=enc, lengths=lens, quals=quals, ambiguous=amb, max_length=ml)


def write_fastq(path, seqs, quals, names=None):
    """4-line FASTQ records from lists of byte strings."""
    with open(path, "wb") as fh:
        for i, (sq, q) in enumerate(zip(seqs, quals)):
            name = names[i] if names else b"r%d" % i
            fh.write(b"@" + name + b"\n" + sq + b"\n+\n" + q + b"\n")


MLP_PARAMS = 64 * 16 + 64 + 64 * 64 + 64 + 64 + 1


def mlp_logits_bf16(params, X):
    """Logits of the learned seam's inference arithmetic (bfloat16 weights/activations)."""
    params = np.ascontiguousarray(params, np.float32)
    X = np.ascontiguousarray(X, np.float32)
    z = np.zeros(len(X), np.float32)
    cpu_lib().carecpu_mlp_logits_bf16(params.ctypes.data, X.ctypes.data, len(X), z.ctypes.data)
    return z


def mlp_loss_grad_f64(params, X, y=None, grad=True):
    """(mean BCE loss, gradient or None, logits) in double precision."""
    params = np.ascontiguousarray(params, np.float64)
    X = np.ascontiguousarray(X, np.float32)
    yy = None if y is None else np.ascontiguousarray(y, np.float32)
    G = np.zeros(MLP_PARAMS, np.float64) if grad else None
    z = np.zeros(len(X), np.float64)
    loss = cpu_lib().carecpu_mlp_loss_grad_f64(params.ctypes.data, X.ctypes.data, _p(yy), len(X), _p(G), z.ctypes.data)
    return loss, G, z


def ref_hash(row, length, k=20, nmaps=48):
    out = np.zeros(nmaps, np.uint64)
    row = np.ascontiguousarray(row, np.uint32)
    ref_lib().ref_hash(row.ctypes.data, int(length), k, nmaps, out.ctypes.data)
    return out


def cpu_hash(row, length, k=20, nmaps=48):
    out = np.zeros(nmaps, np.uint64)
    row = np.ascontiguousarray(row, np.uint32)
    cpu_lib().carecpu_signature(row.ctypes.data, int(length), k, nmaps, out.ctypes.data)
    return out


def ref_align(a, la, c, lc, min_overlap=30, max_error_rate=0.2, ratio=0.3):
    out = np.zeros(5, np.int32)
    a = np.ascontiguousarray(a, np.uint32)
    c = np.ascontiguousarray(c, np.uint32)
    ref_lib().ref_align(a.ctypes.data, la, c.ctypes.data, lc, min_overlap, max_error_rate, ratio, out.ctypes.data)
    return out


def cpu_align(a, la, c, lc, min_overlap=30, max_error_rate=0.2, ratio=0.3):
    out = np.zeros(5, np.int32)
    a = np.ascontiguousarray(a, np.uint32)
    c = np.ascontiguousarray(c, np.uint32)
    cpu_lib().carecpu_align(a.ctypes.data, la, c.ctypes.data, lc, min_overlap, max_error_rate, ratio, out.ctypes.data)
    return out


def _flag_pairs(fn, c1, c2):
    c1 = np.ascontiguousarray(c1, dtype=np.uint32)
    c2 = np.ascontiguousarray(c2, dtype=np.uint32)
    f1 = np.zeros(len(c1), np.uint8)
    f2 = np.zeros(len(c2), np.uint8)
    fn(_p(c1), len(c1), _p(c2), len(c2), _p(f1), _p(f2))
    return f1.astype(bool), f2.astype(bool)


def cpu_flag_pairs(c1, c2):
    """Paired-candidate flags of two mates' sorted candidate lists (oracle restatement)."""
    return _flag_pairs(cpu_lib().carecpu_flag_pairs, c1, c2)


def ref_flag_pairs(c1, c2):
    """The same flags from the reference's findPositionsOfPairedReadIds."""
    return _flag_pairs(ref_lib().ref_flag_pairs, c1, c2)


def parse_records(data: bytes):
    """Splits serialized records into a list of (read_id, flags, payload)."""
    out = []
    p = 0
    mv = memoryview(data)
    while p < len(data):
        rid = int.from_bytes(mv[p:p + 4], "little")
        flags = int.from_bytes(mv[p + 4:p + 8], "little")
        nb = flags & ((1 << 29) - 1)
        out.append((rid, flags, bytes(mv[p + 8:p + 8 + nb])))
        p += 8 + nb
    return out
