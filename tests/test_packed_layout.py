"""Index arithmetic of the packed in-edge copies (CPU only).

The HBM copy and the host tail of a PageRank engine's in-edge store hold ids
packed at b = ceil(log2 nv) bits (id k at bits [k*b, k*b + b)). A pull tile
whose ids all lie in one copy stages the words it spans into its wave's LDS
slot-base array with 16-byte loads (csrc/sg_pma.h, sweep_tile<true>), and the
tail is packed window by window through a bounded HBM stage
(csrc/sg_engine_core.h, refresh_mirror). These tests check the bounds those
kernels rely on for every id width: the staged words fit the LDS array, the
aligned loads stay inside a copy allocated with its 4 words of slack, every
id unpacks from the staged words, and the tail windows start on word
boundaries, fit the stage and cover the tail.
"""
import numpy as np

SWEEP_TILE = 512          # kSweepTile: path positions per tile
LDS_WORDS = SWEEP_TILE + 1  # SweepWave::base[kSweepTile + 1]
SLACK = 4                 # kCopySlack


def packed_words(n, b):
    return (n * b + 31) // 32 + 1


def ids_in_words(words, b):
    return (words - 1) * 32 // b if words > 1 else 0


def stage_span(i0, ne, b):
    """(first aligned word, words staged, bit of id i0 in the staged words)."""
    wa = ((i0 * b) >> 5) & ~3
    nw = ((i0 + ne) * b + 31) // 32 + 1 - wa
    return wa, nw, i0 * b - wa * 32


def pack(ids, b):
    words = np.zeros(packed_words(len(ids), b) + SLACK, dtype=np.uint64)
    for k, v in enumerate(ids):
        bit = k * b
        x = int(v) << (bit & 31)
        words[bit >> 5] |= x & 0xFFFFFFFF
        words[(bit >> 5) + 1] |= x >> 32
    return words


def test_staged_words_fit_lds_and_copy_for_every_width():
    for b in range(1, 32):
        # the word alignment of a tile's first id repeats every 128 ids
        for i0 in range(128):
            for ne in (1, 2, 31, 64, 257, SWEEP_TILE - 1, SWEEP_TILE):
                wa, nw, bit0 = stage_span(i0, ne, b)
                assert wa % 4 == 0 and 0 <= bit0 < 128
                # two 16-byte loads per lane (lanes 0..63: words 4l and 256 + 4l)
                assert nw <= 2 * 64 * 4
                last = max(q for q in range(0, 512, 4) if q < nw)
                assert last + 3 < LDS_WORDS, (b, i0, ne, nw)
                # a tile ending at the copy's last id: the aligned loads stay in
                # the allocation (packed words + slack)
                total = i0 + ne
                assert wa + last + 3 < packed_words(total, b) + SLACK, (b, i0, ne)
                # the last id's second word is staged
                assert ((bit0 + (ne - 1) * b) >> 5) + 1 < nw


def test_staged_ids_unpack():
    rng = np.random.default_rng(11)
    for b in (1, 5, 12, 17, 27, 31):
        total = 3000
        ids = rng.integers(0, 1 << b, size=total, dtype=np.uint64)
        words = pack(ids, b)
        for _ in range(40):
            i0 = int(rng.integers(0, total))
            ne = int(rng.integers(1, min(SWEEP_TILE, total - i0) + 1))
            wa, nw, bit0 = stage_span(i0, ne, b)
            lw = words[wa:wa + nw]
            for p in range(ne):
                bit = bit0 + p * b
                x = (int(lw[(bit >> 5) + 1]) << 32) | int(lw[bit >> 5])
                assert (x >> (bit & 31)) & ((1 << b) - 1) == int(ids[i0 + p])


def test_tail_windows_fit_the_stage_and_cover_the_tail():
    rng = np.random.default_rng(5)
    for b in range(1, 32):
        for _ in range(20):
            rest = int(rng.integers(1, 200000))
            twords = packed_words(rest, b)
            stage = min(max(1024, int(rng.integers(1, 1 << 16))), twords)
            win = max(32, ids_in_words(stage, b) & ~31)
            covered = 0
            for lo in range(0, rest, win):
                hi = min(rest, lo + win)
                used = ((hi - lo) * b + 31) // 32
                assert lo % 32 == 0 and (lo * b) % 32 == 0  # starts on a word of its own
                assert used <= stage, (b, rest, stage, lo)
                assert lo * b // 32 + used <= twords  # inside the host tail
                covered += hi - lo
            assert covered == rest
