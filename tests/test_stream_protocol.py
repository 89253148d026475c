"""A host model of the job stream of k_render_stream (dumbray_amd/csrc/dr_render.hip): the
hand-out of path jobs, the block ids fetched in stream order into a ring, the pixel slots of
the LDS staging ring, the gate on a block ring entry's reuse and the end of the stream.
Waves run one loop iteration at a time in random order, their lanes in lockstep (every lane
reads the LDS state, then the lanes write), paths take random numbers of iterations (a few of
them very many).  Every run must end, with every pixel of every block written exactly once
from all of its samples, and no slot or ring entry reused while a job still needs it.  The
same runs without the gate stall, which the first test shows."""
import random

import pytest

KB = 128                      # pixels per framebuffer block
REFILL_MIN, SHADE_MIN = 8, 32


class Stall(Exception):
    pass


def run(seed, nblocks, spp, nwg, waves=4, maxlen=40, ring=16, gate=True, long_p=0.0, long_len=0,
        stream_samples=2048, max_slots=1024):
    rnd = random.Random(seed)
    nslots = min(stream_samples // spp, max_slots)
    work = [0]
    written = {}
    wgs = []
    for _ in range(nwg):
        wgs.append(dict(owner=list(range(nslots)), remaining=[spp] * nslots, blocks=[0] * ring,
                        px_done=[0] * ring, next_job=0, end_k=1 << 31, staged={},
                        lanes=[[dict(q=-1, live=False, exh=False, blk=0, dur=0) for _ in range(64)]
                               for _ in range(waves)]))

    def finish(wg, L):
        q = L["q"]
        u, s = divmod(q, spp)
        slot = u % nslots
        assert wg["owner"][slot] == u, "slot reused while a job of its previous pixel ran"
        wg["staged"][(slot, s)] = (L["blk"], u % KB)
        wg["remaining"][slot] -= 1
        if wg["remaining"][slot] == 0:
            for k in range(spp):
                assert wg["staged"].pop((slot, k)) == (L["blk"], u % KB)
            key = (L["blk"], u % KB)
            written[key] = written.get(key, 0) + 1
            wg["remaining"][slot] = spp
            wg["px_done"][(u // KB) % ring] += 1
            wg["owner"][slot] = u + nslots
        L["q"], L["live"] = -1, False

    def wave_iter(wg, lanes):
        want = [L["q"] < 0 and not L["exh"] for L in lanes]
        n = sum(want)
        if n and (n >= REFILL_MIN or not any(L["live"] for L in lanes)):
            base = wg["next_job"]
            wg["next_job"] += n
            i = 0
            for L, wnt in zip(lanes, want):
                if wnt:
                    L["q"] = base + i
                    i += 1
        snap = list(wg["blocks"])
        for L in lanes:
            if L["q"] < 0 or L["live"]:
                continue
            q = L["q"]
            u, s = divmod(q, spp)
            k = u // KB
            kk = k % ring
            if k >= wg["end_k"]:
                L["q"], L["exh"] = -1, True
                continue
            if gate and wg["px_done"][kk] < KB * (k // ring):
                continue
            e = snap[kk]
            if e == 0 or (e & 0xff) != (k & 0xff):
                prev = snap[(k - 1) % ring] if k > 0 else 0
                prev_ok = k == 0 or (prev != 0 and (prev & 0xff) == ((k - 1) & 0xff))
                e = 0
                if q == k * KB * spp and prev_ok:
                    bid = work[0]
                    work[0] += 1
                    if bid >= nblocks:
                        wg["end_k"] = min(wg["end_k"], k)
                        L["q"], L["exh"] = -1, True
                    else:
                        e = ((bid + 1) << 8) | (k & 0xff)
                        wg["blocks"][kk] = e
            if e != 0 and (e & 0xff) == (k & 0xff) and wg["owner"][u % nslots] == u:
                L["blk"] = (e >> 8) - 1
                L["live"] = True
                L["dur"] = long_len if rnd.random() < long_p else rnd.randint(0, maxlen)
                if u % KB >= 100 and L["blk"] % 7 == 3:          # pixels outside the image end at once
                    L["dur"] = -1
        for L in lanes:
            if L["live"] and L["dur"] == -1:
                finish(wg, L)
        if not any(L["live"] for L in lanes):
            return "done" if not any(L["q"] >= 0 or not L["exh"] for L in lanes) else "idle"
        for L in lanes:
            if L["live"] and L["dur"] > 0:
                L["dur"] -= 1
        done = [L for L in lanes if L["live"] and L["dur"] == 0]
        if done and (len(done) >= SHADE_MIN or not any(L["live"] and L["dur"] > 0 for L in lanes)):
            for L in done:
                finish(wg, L)
        return "busy"

    active = [(g, w) for g in range(nwg) for w in range(waves)]
    idle = 0
    while active:
        i = rnd.randrange(len(active))
        g, w = active[i]
        r = wave_iter(wgs[g], wgs[g]["lanes"][w])
        if r == "done":
            active.pop(i)
            idle = 0
        elif r == "idle":
            idle += 1
            if idle > 20000:              # every remaining wave waits: nothing can progress
                raise Stall()
        else:
            idle = 0
    assert all(not g["staged"] for g in wgs)
    assert set(written) == {(b, lp) for b in range(nblocks) for lp in range(KB)}
    assert all(v == 1 for v in written.values())


@pytest.mark.parametrize("seed,nblocks,spp,nwg", [(1, 40, 16, 3), (2, 5, 16, 8), (3, 60, 1, 4), (4, 33, 4, 3),
                                                  (5, 9, 64, 2), (6, 1, 256, 3), (7, 100, 2, 5), (8, 17, 16, 1)])
def test_stream_writes_every_pixel_once(seed, nblocks, spp, nwg):
    run(seed, nblocks, spp, nwg)


@pytest.mark.parametrize("ring", [2, 4, 16])
@pytest.mark.parametrize("seed", range(3))
def test_stream_with_long_paths(ring, seed):
    rnd = random.Random(1000 * ring + seed)
    run(1000 * ring + seed, rnd.randint(1, 40), rnd.choice([1, 2, 4, 16]), rnd.randint(1, 3), ring=ring,
        waves=rnd.choice([1, 2, 4]), long_p=0.003, long_len=rnd.choice([500, 2000]))


def test_ring_reuse_needs_the_gate():
    """Without the gate a lane can wait on a ring entry that a later block has taken over."""
    with pytest.raises(Stall):
        run(11, 120, 1, 2, long_p=0.002, long_len=20000, gate=False)
