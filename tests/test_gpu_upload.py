"""Large read-storage uploads (care_hip_set_reads).

Arrays of 64 MiB or more go to the device through the library's upload threads (slices
packed into pinned double buffers, one stream per thread); quality rows whose pitch is
not a multiple of 16 are padded to the device pitch while they are packed. Bar: the
storage read back through care_hip_inspect_reads equals the host arrays, whole.
"""
import numpy as np
import pytest

from rabbitqcplus_amd import engine, reads

pytestmark = pytest.mark.gpu


def _random_storage(n, max_len, qual_width, seed):
    rng = np.random.default_rng(seed)
    lengths = rng.integers(max_len // 2, max_len + 1, n).astype(np.int32)
    lengths[0] = max_len
    pitch = reads.ints_per_row(max_len)
    words = rng.integers(0, 1 << 32, (n, pitch), dtype=np.uint64)
    # bases past a read's end are zero: base i sits at bits 31 - 2 (i % 16) of word i // 16
    nb = np.clip(lengths[:, None].astype(np.int64) - 16 * np.arange(pitch)[None, :], 0, 16).astype(np.uint64)
    keep = (((np.uint64(1) << (2 * nb)) - np.uint64(1)) << (np.uint64(32) - 2 * nb)) & np.uint64(0xFFFFFFFF)
    enc = (words & keep).astype(np.uint32)
    quals = rng.integers(33, 75, (n, qual_width), dtype=np.uint8)
    quals[np.arange(qual_width)[None, :] >= lengths[:, None]] = 0
    ambiguous = (rng.random(n) < 0.01).astype(np.uint8)
    return reads.ReadStorage(enc=enc, lengths=lengths, quals=quals, ambiguous=ambiguous)


@pytest.mark.parametrize("qual_width", [250, 256])
def test_large_upload_reads_back_equal(qual_width):
    """1.2M reads of up to 250 bases: 77 MB of rows and 300-307 MB of quality rows (a
    250-byte pitch padded to 256 on the device, or already 256) take the threaded upload."""
    n, max_len = 1_200_000, 250
    rs = _random_storage(n, max_len, qual_width, seed=qual_width)
    assert rs.enc.nbytes >= 64 << 20 and rs.quals.nbytes >= 64 << 20
    print(f"[upload] reads={n} rows={rs.enc.nbytes >> 20} MiB quals={rs.quals.nbytes >> 20} MiB", flush=True)
    eng = engine.CareEngine(0, use_quality_scores=1)
    try:
        eng.set_reads(rs)
        assert eng.read_layout() == (n, max_len, rs.pitch_ints)
        got = eng.inspect_reads()
        # a second upload into the same context (buffers and upload threads reused)
        rs2 = _random_storage(n, max_len, qual_width, seed=qual_width + 1)
        eng.set_reads(rs2)
        got2 = eng.inspect_reads()
    finally:
        eng.close()
    for want, g in ((rs, got), (rs2, got2)):
        assert np.array_equal(g["lengths"], want.lengths)
        assert np.array_equal(g["enc"], want.enc)
        assert np.array_equal(g["quals"], want.quals[:, :max_len])
        assert np.array_equal(g["ambiguous"], want.ambiguous)
