"""CPU model of the line-aligned strided pass (gpu_fft_amd/csrc/gfft_lines.h).

The kernel reads each row of a plane as aligned CW-element chunks and keeps
the part of a chunk that belongs to the next tile in registers; stores write
whole chunks, the first positions from the previous tile's outputs.  This
model runs the kernel's index arithmetic (tile walk, sigma, chunk biases,
register rotation, masked boundary stores) for every row and lane of every
workgroup, in each workgroup's program order, on small geometries, and
checks:

- every tile sees exactly its own columns as inputs (kz = CW*j + c for the
  lane's column c = (l - sigma) mod CW), zeros past the end of a row;
- every element of every row is written exactly once, by the workgroup that
  owns its tile, with that tile's output for it, and nothing outside the
  rows' [0, inner) columns is written;
- a workgroup reads every element it owns (the load whose value it uses)
  before it writes it, so in-place execution is safe.
"""
import random

import pytest

OOR = 0x80000000


def walk_ranges(ntiles, G):
    """tile_walk<0>: workgroup w owns [w*T/G, (w+1)*T/G)."""
    out = []
    for w in range(G):
        a, b = w * ntiles // G, (w + 1) * ntiles // G
        out.append((a, b - a))
    return out


def simulate(N, inner, planes, CW, rlo, outer, base_mod, G, rev, mistake=None):
    es = 1  # element units
    tpp = -(-inner // CW)
    ntiles = planes * tpp
    plane_bytes = (N - 1) * rlo + inner

    def sigma_of(plane, n):
        return (base_mod + plane * outer + n * rlo) % CW

    def bias(m, l, sigma):
        kz = CW * m + l - sigma
        return kz * es if 0 <= kz < inner else OOR

    # memory: element address (plane, offset) -> value; value = ("in", plane, n, kz) or output tags
    writes = {}      # (plane, off) -> list of (wg, value)
    read_used = {}   # (plane, n, kz) -> (wg, time of the load)
    write_time = {}  # (plane, off) -> list of (wg, time)
    owner = {}
    for w, (first, count) in enumerate(walk_ranges(ntiles, G)):
        for k in range(count):
            owner[(ntiles - 1 - (first + k)) if rev else first + k] = w

    def mem_load(plane, off, wg, time):
        if off >= plane_bytes:  # range check (OOR included)
            return None
        n, kz = divmod(off, rlo)
        assert kz < inner or off >= plane_bytes, "load inside a row gap"
        return ("in", plane, n, kz, wg, time)

    for w, (first, count) in enumerate(walk_ranges(ntiles, G)):
        if count == 0:
            continue
        time = 0
        last = ntiles - 1

        def at(k):
            t = first + k
            return last - t if rev else t

        fwd = not rev
        tile = at(0)
        plane, j = divmod(tile, tpp)
        # per (row, lane) register state
        def load(plane_, m, n, l, sig):
            nonlocal time
            time += 1
            b = bias(m, l, sig)
            off = n * rlo + b if b != OOR else OOR + n * rlo
            return mem_load(plane_, off, w, time)

        older, newer = {}, {}
        # uninitialised registers before the first tile
        carry = {(n, l): ("garbage",) for n in range(N) for l in range(CW)}
        for n in range(N):
            for l in range(CW):
                s = sigma_of(plane, n)
                older[n, l] = load(plane, j if fwd else j + 1, n, l, s)
                newer[n, l] = load(plane, j + 1 if fwd else j, n, l, s)
        have_prev = False
        for k in range(count):
            cur = {}
            for n in range(N):
                for l in range(CW):
                    s = sigma_of(plane, n)
                    take_newer = (l < s) == fwd
                    if mistake == "select":
                        take_newer = not take_newer
                    cur[n, l] = newer[n, l] if take_newer else older[n, l]
                    col = (l - s) % CW
                    kz = CW * j + col
                    v = cur[n, l]
                    if kz < inner:
                        assert v is not None and v[1:4] == (plane, n, kz), (v, plane, n, kz)
                        key = (plane, n, kz)
                        assert key not in read_used
                        read_used[key] = (w, v[5])
                    else:
                        assert v is None, v
            next_same = False
            plane2, j2 = plane, j
            if k + 1 < count:
                plane2, j2 = divmod(at(k + 1), tpp)
                next_same = plane2 == plane
                for n in range(N):
                    for l in range(CW):
                        if next_same:
                            older[n, l] = newer[n, l]
                            newer[n, l] = load(plane, j2 + 1 if fwd else j2, n, l, sigma_of(plane, n))
                        else:
                            s2 = sigma_of(plane2, n)
                            older[n, l] = load(plane2, j2 if fwd else j2 + 1, n, l, s2)
                            newer[n, l] = load(plane2, j2 + 1 if fwd else j2, n, l, s2)
            # outputs of this tile: value tagged with (plane, n, kz of the lane's column)
            out = {}
            for n in range(N):
                for l in range(CW):
                    s = sigma_of(plane, n)
                    out[n, l] = ("out", plane, n, CW * j + (l - s) % CW)

            def store(m, n, l, sig, value):
                nonlocal time
                time += 1
                b = bias(m, l, sig)
                off = n * rlo + b if b != OOR else OOR + n * rlo
                if off >= plane_bytes:
                    return
                writes.setdefault((plane, off), []).append((w, value))
                write_time.setdefault((plane, off), []).append((w, time))

            for n in range(N):
                for l in range(CW):
                    s = sigma_of(plane, n)
                    mine = (l >= s) == fwd
                    if mine or have_prev or mistake == "carry":
                        store(j if fwd else j + 1, n, l, s, out[n, l] if mine else carry[n, l])
            if not next_same and mistake != "flush":
                for n in range(N):
                    for l in range(CW):
                        s = sigma_of(plane, n)
                        if (l < s) == fwd:
                            store(j + 1 if fwd else j, n, l, s, out[n, l])
            carry = out
            have_prev = next_same
            plane, j = plane2, j2

    # every element written once, by its owner, with its own output, after its read
    for p in range(planes):
        for n in range(N):
            for kz in range(inner):
                off = n * rlo + kz
                ws = writes.get((p, off), [])
                assert len(ws) == 1, (p, n, kz, ws)
                wg, val = ws[0]
                assert val == ("out", p, n, kz), (val, p, n, kz)
                tile = p * tpp + kz // CW
                assert wg == owner[tile], (p, n, kz, wg, owner[tile])
                rw, rt = read_used[(p, n, kz)]
                assert rw == wg
                assert write_time[(p, off)][0][1] > rt
    # nothing else written (row gaps when rlo > inner, other planes' space)
    total = sum(len(v) for v in writes.values())
    assert total == planes * N * inner


CASES = [
    # N, inner, planes, CW, rlo, outer, base_mod, G
    (8, 17, 3, 4, 17, 8 * 17, 0, 2),      # y pass, K = 4*4 + 1
    (8, 17, 3, 4, 17, 8 * 17, 3, 5),      # misaligned base, ranges cut inside planes
    (12, 9, 4, 4, 9, 12 * 9, 1, 7),       # K = 9, N not a power of two
    (8, 16, 2, 4, 16, 8 * 16, 0, 3),      # aligned rows: sigma 0 everywhere
    (8, 10, 2, 4, 13, 8 * 13 + 5, 2, 3),  # row gap (rlo > inner), plane gap
    (8, 24, 1, 4, 24, 0, 0, 4),           # x-pass geometry: one plane
    (4, 33, 5, 8, 33, 4 * 33, 5, 6),      # CW = 8
    (8, 3, 3, 4, 3, 8 * 3, 1, 4),         # inner < CW
]


@pytest.mark.parametrize("rev", [False, True])
@pytest.mark.parametrize("case", CASES)
def test_lines_walk_ownership(case, rev):
    N, inner, planes, CW, rlo, outer, base_mod, G = case
    simulate(N, inner, planes, CW, rlo, outer, base_mod, G, rev)


def test_lines_walk_random_geometries():
    rng = random.Random(7)
    for _ in range(40):
        CW = rng.choice([2, 4, 8])
        N = CW * rng.randint(1, 3)
        inner = rng.randint(1, 5 * CW)
        rlo = inner + rng.choice([0, 0, 0, 1, CW - 1])
        planes = rng.randint(1, 4)
        outer = (N - 1) * rlo + inner + rng.choice([0, 0, 2])
        if rng.random() < 0.5:
            outer = N * rlo
        base_mod = rng.randrange(CW)
        G = rng.randint(1, 9)
        simulate(N, inner, planes, CW, rlo, outer, base_mod, G, rng.random() < 0.5)


@pytest.mark.parametrize("mistake", ["select", "carry", "flush"])
def test_planted_mistakes_are_caught(mistake):
    """Each of these kernel mistakes fails the model: the inverted chunk
    choice (wrong inputs), storing carried positions without a previous tile
    (a workgroup writes its neighbour's or another row's elements), and no
    flush at the end of a run (elements never written)."""
    with pytest.raises(AssertionError):
        simulate(8, 17, 3, 4, 17, 8 * 17, 3, 5, False, mistake=mistake)
