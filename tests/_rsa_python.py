"""A second restatement of the reference's draw rounds, in plain Python, written
independently of the C++ oracle -- test infrastructure.

It replays the draws the oracle recorded (candidates in shot order with their
priorities and radii) through the reference's cell storages and conflict rounds
and returns, per draw, which candidates end Validated, and at the end the
extraction order.  The pieces that carry order, and that the oracle restates
too, are spelled out here from the reference's files:

- cells: growable per-cell lists, appended in candidate order (build_grid,
  operator_cells_and_data_storage.ixx:153-159); removal of a tag scans a cell
  from the back and replaces a removed entry by the cell's last one
  (rsa_data_storage.ixx:108-116, 305-315);
- periodic images: a real sphere of an edge cell whose centre lies in the
  closed band of width 2 * domain radius is imaged, shifted by -+L, into the
  ghost cell next to the opposite face; a ghost cell's images follow the
  in-cell order of the one real cell they come from (fill_ghost walks the edge
  cells and their spheres in order, operator_ghost.ixx:5-27;
  rsa_ghost_area.ixx:22-30);
- rounds (operator_algorithm.ixx:383-473): validated filter, then the priority
  sweep where a candidate's first colliding, not-Skip, strictly-smaller-priority
  sphere in stencil then in-cell order decides Conflict (an image, or a
  Conflict sphere) or Skip (operator_cells_and_data_storage.ixx:16-52, 79-143);
  images keep the tag they were put with for the whole sweep; outreach draws
  keep the round's Undecided candidates whose priority is not above the pivot
  of median_of_medians (retain_only_first_elements, operator_algorithm.ixx:475-519,
  median_of_medians.ixx:8-178, transcribed below as written); each round
  compares with the draw-start class count (a by-value parameter at :383).

The stencil order is libstdc++'s std::sort of the offsets by squared norm
(rsa_grid_traversal.ixx:30-66), which the caller passes in (the oracle's table,
itself checked to be a norm sort).
"""
import math

import numpy as np

VALIDATED, UNDECIDED, CONFLICT, SKIP = 0, 1, 2, 3

EPS = 2.220446049250313e-16  # std::numeric_limits<double>::epsilon()


def to_u64(x):
    """double -> uint64_t as g++ emits it on x86-64 (cvttsd2si: truncation toward zero,
    negative results wrapping modulo 2^64); std::less<uint64_t> converts the pivot so."""
    if x != x:
        return 1 << 63
    if x < 9223372036854775808.0:
        k = math.trunc(x)
        if k < -(1 << 63):
            k = -(1 << 63)
        return k % (1 << 64)
    return (math.trunc(x - 9223372036854775808.0) % (1 << 64)) ^ (1 << 63)


def _upper_bound(seq, value, comp):
    """libstdc++ std::upper_bound: first position whose element e has comp(value, e)."""
    first, length = 0, len(seq)
    while length > 0:
        half = length >> 1
        middle = first + half
        if comp(value, seq[middle]):
            length = half
        else:
            first = middle + 1
            length = length - half - 1
    return first


def reference_pivot(values, nb_first):
    """median_of_medians::find_pivot_for_first_elements on one rank with std::less<uint64_t>
    (median_of_medians.ixx:8-123, find_pivot :123-152, nb_smaller_than_pivot :154-178),
    transcribed as written: the candidates' dichotomy starts from the lower candidate and
    steps by half the previous step, upwards while too many values lie below the pivot and
    downwards while too few (so it can leave fewer than nb_first values below)."""
    gsize = len(values)
    if nb_first > gsize:
        raise RuntimeError("Impossible to find the pivot for N first elements if N is smaller than the size of the array")
    if nb_first <= 0:
        raise RuntimeError("Impossible to find the pivot for N first elements if N  <= 0")
    if gsize == 0:
        raise RuntimeError("Impossible to find the pivot for N first elements if no element")
    pct = nb_first / (1.0 * gsize)
    v = sorted(int(x) for x in values)
    size_pct = float(len(v)) * pct
    piv = int(size_pct) - 1
    mn, mx = piv, piv + 1
    if abs(size_pct - 1 - piv) < 1e-8:
        mn -= 1
    if mn < 0:
        mn = 0
    if abs(size_pct - 1 - piv - 1) < 1e-8:
        mx += 1
    if mx > len(v) - 1:
        mx = len(v) - 1
    first = float(v[mn]) - EPS
    secnd = float(v[mx]) + EPS
    cand = [first, secnd]
    if to_u64(cand[1]) < to_u64(cand[0]):  # std::sort of two elements (insertion step)
        cand = [cand[1], cand[0]]

    def below(p):  # upper_bound(sorted, p, less<uint64_t>) - begin
        return _upper_bound(v, p, lambda a, b: to_u64(a) < b)

    it = _upper_bound(cand, nb_first, lambda nb, p: nb <= below(p))
    if it == len(cand):
        raise RuntimeError("Unexpected - 0")
    pmax = cand[it]
    pivot = pmax
    if below(pivot) != nb_first:
        if it == 0:
            raise RuntimeError("Unexpected - 1")
        pmin = cand[it - 1]
        add = 0.5 * (pmax - pmin)
        for _ in range(20):
            pivot = pmin + add
            n = below(pivot)
            if n == nb_first:
                break
            if n > nb_first:
                add = 0.5 * abs(add)
            if n < nb_first:
                add = -0.5 * abs(add)
    return pivot


class Grid:
    """rsa_grid_traversal::define (rsa_grid_traversal.ixx:11-28) with ghost layer gl."""

    def __init__(self, dim, inf, sup, rad, stencil, gl=1):
        self.dim, self.gl = dim, gl
        self.inf = [float(v) for v in inf]
        self.sup = [float(v) for v in sup]
        self.n, self.cl, self.inv, self.shift = [], [], [], []
        for d in range(dim):
            length = rad * 2
            width = self.sup[d] - self.inf[d]
            nin = int(width / length)
            cl = width / nin
            self.cl.append(cl)
            self.inv.append(1.0 / cl)
            self.shift.append(self.inf[d] - gl * cl)
            self.n.append(nin + 2 * gl)
        # offsets in stencil order: entry i of the 3^D table encodes (i // 3^d) % 3 - 1
        self.offsets = [[(int(i) // 3 ** d) % 3 - 1 for d in range(dim)] for i in stencil]

    def coords(self, p):
        """compute_cell_idx's per-dimension terms (rsa_grid_traversal.ixx:373-382)."""
        return tuple(int((p[d] - self.shift[d]) * self.inv[d]) for d in range(self.dim))

    def index(self, c):
        r = 0
        for d in range(self.dim - 1, -1, -1):
            r = r * self.n[d] + c[d]
        return r

    def is_real(self, c):
        return all(self.gl <= c[d] <= self.n[d] - 1 - self.gl for d in range(self.dim))


def collide(r1, p1, r2, p2):
    """is_there_collision (operator_cells_and_data_storage.ixx:5-14), same summation order."""
    s = 0.0
    for a, b in zip(p1, p2):
        dr = a - b
        s += dr * dr
    return s < (r1 + r2) * (r1 + r2)


class Sim:
    """Real cells keyed by coordinates; each holds a list of entries [pos, rad, prio, tag, key]."""

    def __init__(self, grid, drad):
        self.g = grid
        self.drad = drad
        self.cells = {}
        self.L = [grid.sup[d] - grid.inf[d] for d in range(grid.dim)]
        self.pivots = []  # (sorted Undecided priorities, maximal, pivot) of each outreach retention

    def neighbour(self, c, off):
        """(source real cell coords, per-dimension image side) of the stencil neighbour c + off:
        side -1 is the lower ghost layer (images of the upper band, shifted by -L), +1 the upper."""
        g = self.g
        src, side = [], []
        for d in range(g.dim):
            k = c[d] + off[d]
            inner = g.n[d] - 2 * g.gl
            s = 0
            if k < g.gl:
                k += inner
                s = -1
            elif k > g.n[d] - 1 - g.gl:
                k -= inner
                s = 1
            src.append(k)
            side.append(s)
        return tuple(src), side

    def image(self, p, side):
        """The entry's image for a ghost neighbour, or None outside the area's closed band."""
        q = list(p)
        for d, s in enumerate(side):
            if s == -1:  # upper band [sup - 2 rad, sup], shifted by -L
                if p[d] < self.g.sup[d] - 2 * self.drad or p[d] > self.g.sup[d]:
                    return None
                q[d] = p[d] - self.L[d]
            elif s == 1:  # lower band [inf, inf + 2 rad], shifted by +L
                if p[d] < self.g.inf[d] or p[d] > self.g.inf[d] + 2 * self.drad:
                    return None
                q[d] = p[d] + self.L[d]
        return q

    def stencil_entries(self, c):
        """Entries seen from real cell c in detection order: (entry, position seen, is_image)."""
        for off in self.g.offsets:
            src, side = self.neighbour(c, off)
            ghost = any(side)
            for e in self.cells.get(src, []):
                if not ghost:
                    yield e, e[0], False
                else:
                    q = self.image(e[0], side)
                    if q is not None:
                        yield e, q, True

    def remove(self, c, pred):
        """remove_tagged_spheres: from the back, the removed entry replaced by the last one."""
        lst = self.cells.get(c, [])
        for i in range(len(lst) - 1, -1, -1):
            if pred(lst[i]):
                lst[i] = lst[-1]
                lst.pop()

    def draw(self, pos, rad, prio, maximal, may_outreach):
        """One draw (single_draw + add_to_sample rounds); returns the accepted flags."""
        n = len(rad)
        touched = []
        for i in range(n):
            c = self.g.coords(pos[i])
            if not self.g.is_real(c):
                continue  # rounded into a ghost cell: never placed (the reference drops it)
            self.cells.setdefault(c, []).append([list(pos[i]), float(rad[i]), int(prio[i]), UNDECIDED, i])
            if c not in touched:
                touched.append(c)
        while True:
            # remove_collisions_with_validated_spheres, cells from the back (:54-76)
            for c in touched:
                for e in reversed(self.cells[c]):
                    if e[3] != UNDECIDED:
                        continue
                    for f, q, img in self.stencil_entries(c):
                        if f[3] == VALIDATED and collide(e[1], e[0], f[1], q):
                            e[3] = SKIP
                            break
                self.remove(c, lambda e: e[3] == SKIP)
            # the sweep: images are put with the tag Undecided and keep it
            und = [(e[2], c, e) for c in touched for e in self.cells[c] if e[3] == UNDECIDED]
            und.sort(key=lambda t: t[0])
            frozen = {id(e): e[3] for c in touched for e in self.cells[c]}
            for pi, c, x in und:
                for f, q, img in self.stencil_entries(c):
                    tag = UNDECIDED if img else f[3]
                    if img and frozen.get(id(f), f[3]) != UNDECIDED:
                        continue  # only Undecided spheres are imaged at the round's start
                    if tag == SKIP or not (f[2] < pi) or not collide(x[1], x[0], f[1], q):
                        continue
                    x[3] = CONFLICT if (img or tag == CONFLICT) else SKIP
                    break
            undecided = [e for c in touched for e in self.cells[c] if e[3] == UNDECIDED]
            total = len(undecided) if may_outreach else 0
            if may_outreach and total > maximal:
                # retain_only_first_elements: Undecided above the pivot and every Conflict become Skip
                pivot = reference_pivot([e[2] for e in undecided], maximal)
                self.pivots.append((sorted(e[2] for e in undecided), maximal, pivot))
                for c in touched:
                    for e in self.cells[c]:
                        if e[3] == UNDECIDED and float(e[2]) > pivot:
                            e[3] = SKIP
                        elif e[3] == CONFLICT:
                            e[3] = SKIP
            for c in touched:
                self.remove(c, lambda e: e[3] == SKIP)
                for e in self.cells[c]:
                    if e[3] == UNDECIDED:
                        e[3] = VALIDATED
                    elif e[3] == CONFLICT:
                        e[3] = UNDECIDED
            if may_outreach and total > maximal:
                break
            if maximal - total == 0:
                for c in touched:
                    self.remove(c, lambda e: e[3] != VALIDATED)
                break
            if not any(e[3] != VALIDATED for c in touched for e in self.cells[c]):
                break
        acc = np.zeros(n, dtype=bool)
        for c in touched:
            for e in self.cells[c]:
                if e[3] == VALIDATED and e[4] is not None:
                    acc[e[4]] = True
                    e[4] = None  # placed: no longer a candidate of later draws
        return acc

    def extraction(self):
        """rsa_grid::extract_spheres (rsa_grid.ixx:73-98): real cells by index, in-cell order."""
        out_p, out_r = [], []
        for c in sorted(self.cells, key=self.g.index):
            for e in self.cells[c]:
                out_p.append(e[0])
                out_r.append(e[1])
        return np.array(out_p), np.array(out_r)
