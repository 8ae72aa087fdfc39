"""Host emulation of the batched engine's segmented backward mapping (rasterSegmentedBatchedKernel
with forwardBlock's checkpoints, csrc/kernels_raster.hip), with bounds checks on every index.

The batched engine keeps this path switched off: its first GPU run did not finish and the cause is
not established. This emulation restates, with the kernels' own integer formulas, everything the
path computes from sizes and counts -- global tile ids over several views, the owning view of a
tile (itemOfBlock), the owning tile of a work item (tileOfItem over the exclusive scan of segment
counts), the checkpoint slots the forward writes (begin / kBatch + tile + k) and the ones the
backward reads, the records a segment stages -- for views of uneven sizes, empty and long tile
lists, pixels that stop early, every LDS batch (64, 128, 256) and pair lists cut at the capacity.
It finds every index inside its buffer, every checkpoint read written by the forward and every
(tile, segment) work item visited once, so the claim loop ends after sum(segCount) items.
"""
import numpy as np
import pytest

THREADS = 256
STEP = 64


def _tile_of_item(seg_start, tiles, w):
    lo, hi = 0, tiles - 1
    while lo < hi:
        mid = (lo + hi + 1) >> 1
        if seg_start[mid] <= w:
            lo = mid
        else:
            hi = mid - 1
    return lo


def _item_of_block(block_base, b):
    lo, hi = 0, len(block_base) - 1
    while lo < hi:
        mid = (lo + hi + 1) >> 1
        if block_base[mid] <= b:
            lo = mid
        else:
            hi = mid - 1
    return lo


def _views(rng, count):
    sizes = []
    for _ in range(count):
        w = int(rng.integers(17, 200))
        sizes.append((w, int(rng.integers(9, 150))))
    return sizes


@pytest.mark.parametrize("k_batch", [64, 128, 256])
@pytest.mark.parametrize("seed,cut", [(1, 1.0), (2, 0.7), (3, 1.0)])
def test_batched_segment_mapping_stays_in_bounds(k_batch, seed, cut):
    rng = np.random.default_rng(seed * 10 + k_batch)
    views = _views(rng, 5)
    blocks_x = [(w + 15) // 16 for w, _ in views]
    blocks_y = [(h + 15) // 16 for _, h in views]
    counts = [bx * by for bx, by in zip(blocks_x, blocks_y)]
    base = np.concatenate([[0], np.cumsum(counts)[:-1]]).astype(np.int64)  # blockBase == tileBase
    tiles = int(sum(counts))
    # list lengths: empty, short, and long (several segments), then cut at the pair capacity
    kind = rng.integers(0, 4, tiles)
    length = np.where(kind == 0, 0, np.where(kind == 1, rng.integers(1, 64, tiles),
                                             rng.integers(64, 1400, tiles)))
    total_pairs = int(length.sum())
    pair_cap = max(int(total_pairs * cut), 1)
    kept = length.copy()
    over = total_pairs - min(total_pairs, pair_cap)
    for t in rng.permutation(tiles):  # pairs past the capacity are dropped from some lists
        if over <= 0:
            break
        d = min(int(kept[t]), over)
        kept[t] -= d
        over -= d
    offsets = np.concatenate([[0], np.cumsum(kept)]).astype(np.int64)
    assert offsets[-1] <= pair_cap
    ckpt_cap = (pair_cap // k_batch + tiles + 1) * THREADS  # BatchEngine::ensureSegments
    written = {}
    n_contrib = {}
    seg_count = np.zeros(tiles, np.int64)
    # forward: per tile, per pixel of a 16x16 block
    for v, (W, H) in enumerate(views):
        for local in range(counts[v]):
            bx, by = local % blocks_x[v], local // blocks_x[v]
            tile = int(base[v] + by * blocks_x[v] + bx)
            begin, end = int(offsets[tile]), int(offsets[tile + 1])
            L = end - begin
            tid = np.arange(THREADS)
            wave, lane = tid >> 6, tid & 63
            px = bx * 16 + (wave & 1) * 8 + (lane & 7)
            py = by * 16 + (wave >> 1) * 8 + (lane >> 3)
            inside = (px < W) & (py < H)
            stopped = rng.random(THREADS) < 0.5
            stop_at = rng.integers(0, L + 1, THREADS) if L else np.zeros(THREADS, np.int64)
            nc = np.where(stopped, rng.integers(0, stop_at + 1), rng.integers(0, L + 1, THREADS))
            nc = np.where(inside, nc, 0)
            assert (nc <= L).all()
            n_contrib[tile] = nc
            slot0 = begin // k_batch + tile
            for w4 in range(4):
                lanes = wave == w4
                # steps the wavefront walks: until every lane is done (outside lanes are done)
                ends = np.where(~inside[lanes], 0,
                                np.where(stopped[lanes], stop_at[lanes] // STEP + 1, (L + STEP - 1) // STEP))
                steps = min((L + STEP - 1) // STEP, int(ends.max()))
                batches = 0
                for s in range(steps):
                    b = s * STEP
                    if b % k_batch == 0:
                        if b > 0:
                            for t in tid[lanes & inside]:
                                idx = (slot0 + batches) * THREADS + int(t)
                                assert 0 <= idx < ckpt_cap
                                written[(tile, int(t), batches)] = idx
                        batches += 1
            seg_count[tile] = (int(nc.max()) + k_batch - 1) // k_batch
    # backward: the claim loop over (tile, segment) work items
    seg_start = np.concatenate([[0], np.cumsum(seg_count)[:-1]]).astype(np.int64)
    total = int(seg_count.sum())
    visited = set()
    for w in range(total):
        tile = _tile_of_item(seg_start, tiles, w)
        v = _item_of_block(base, tile)
        local = tile - int(base[v])
        assert 0 <= local < counts[v]
        bx, by = local % blocks_x[v], local // blocks_x[v]
        assert int(base[v] + by * blocks_x[v] + bx) == tile  # backwardSegment's tile id
        seg = w - int(seg_start[tile])
        assert 0 <= seg < seg_count[tile]
        begin, end = int(offsets[tile]), int(offsets[tile + 1])
        lo, hi = seg * k_batch, seg * k_batch + k_batch
        seg_len = min(k_batch, end - begin - lo)
        nc = n_contrib[tile]
        max_n = int(np.minimum(nc, hi).max())  # the largest jEnd = min(nContrib, hi)
        assert max_n > lo  # a counted segment has a pixel that reaches it
        assert 0 < max_n - lo <= seg_len  # staged records exist for every entry walked
        assert begin + lo + seg_len <= offsets[-1] <= pair_cap
        for t in np.nonzero(nc > hi)[0]:  # pixels starting from a checkpoint
            idx = (begin // k_batch + tile + seg + 1) * THREADS + int(t)
            assert 0 <= idx < ckpt_cap
            assert written.get((tile, int(t), seg + 1)) == idx
        assert (tile, seg) not in visited
        visited.add((tile, seg))
    assert len(visited) == total
