"""MI355X-native single-rank RSA (random sequential adsorption) with the
operator API of MarcJos/RSA_MPI.

This module mirrors the reference's Python module ``rsa_mpi_py``
(Interface_python/Interface.cxx:40-120): per dimension a ``sphere``,
``rsa_domain``, ``radius_generator``, ``rsa_algo`` class and a ``throw_spheres``
function (``rsa_domain_3D``, ``rsa_algo_3D``, ...).  All work runs in
librsa_mpi_amd.so (hand-written gfx950 HIP kernels behind a C-ABI); this file
only marshals arguments.
"""
import ctypes

import numpy as np

from . import _lib
from . import tuning as _tuning
from .tuning import autotune, autotune_batch  # noqa: F401
from ._lib import Stats, check, dptr, f64, lib, mt_words  # noqa: F401

__all__ = ["Domain", "RadiusGenerator", "LawRadiusGenerator", "Algo", "Stats", "mt_words", "uniform_generate",
           "throw_spheres", "batch_run", "batch_queue", "write_binary", "read_binary", "set_nb_threads", "learn_reset",
           "learn_train", "learn_samples", "learn_params", "learn_gradcheck", "learn_forward"]


def set_nb_threads(n):
    """rsa_mpi_py.set_nb_threads (Interface.cxx:24-35): the GPU path has no OpenMP pool."""
    return None


class Sphere:
    """Sphere<DIM> (auxiliary/Geometry/Sphere.hxx:18-32): center, radius, phase."""

    def __init__(self, center=None, radius=0.0, phase=0):
        self.center = list(center) if center is not None else []
        self.radius = float(radius)
        self.phase = int(phase)

    def __repr__(self):
        return f"Sphere(center={self.center}, radius={self.radius}, phase={self.phase})"


class Domain:
    """rsa_domain<DIM> on one GPU (rsa_domain.hxx:53-93): periodic box [inf, sup]."""

    def __init__(self, inf, sup, rad, ghost_layer=1, tuning="auto"):
        """tuning: "auto" applies the thresholds autotune() stored for this dimension and
        box-to-radius ratio (if any), None keeps the defaults, a dict sets them."""
        self.inf = [float(v) for v in inf]
        self.sup = [float(v) for v in sup]
        self.dim = len(self.inf)
        if len(self.sup) != self.dim:
            raise ValueError("Incoherent dimensions!")
        self.rad = float(rad)
        self._ctx = ctypes.c_void_p()
        a, b = f64(self.inf), f64(self.sup)
        check(lib().rsa_gpu_create(self.dim, dptr(a), dptr(b), int(ghost_layer), self.rad, ctypes.byref(self._ctx)))
        if tuning == "auto":
            side = min(b_ - a_ for a_, b_ in zip(self.inf, self.sup))
            tuning = _tuning.lookup(self.dim, side, self.rad) if self.rad > 0 else None
        if tuning:
            self.set_tuning(**{k: tuning[k] for k in ("group_min", "pack_min") if k in tuning})

    def set_tuning(self, group_min=None, pack_min=None):
        """The refinement thresholds (rsa_gpu_set_tuning): they choose how voxel lists are
        walked, never the result."""
        t = self.tuning()
        t = _lib.Tuning(int(group_min if group_min is not None else t["group_min"]),
                        int(pack_min if pack_min is not None else t["pack_min"]))
        check(lib().rsa_gpu_set_tuning(self._ctx, ctypes.byref(t)))

    def tuning(self):
        t = _lib.Tuning()
        check(lib().rsa_gpu_get_tuning(self._ctx, ctypes.byref(t)))
        return {"group_min": t.group_min, "pack_min": t.pack_min}

    def __del__(self):
        ctx = getattr(self, "_ctx", None)
        if ctx is not None and ctx.value:
            lib().rsa_gpu_destroy(ctx)
            self._ctx = None

    def get_total_volume(self):  # rsa_domain.ixx:166-169
        v = 1.0
        for a, b in zip(self.inf, self.sup):
            v *= (b - a)
        return v

    def domain_log(self):
        print(" ==== log: domain [0] ==== ")
        print(" == m_inf : " + " ".join(str(v) for v in self.inf))
        print(" == m_sup : " + " ".join(str(v) for v in self.sup))

    def number_of_spheres(self):
        n = ctypes.c_uint64()
        check(lib().rsa_gpu_count(self._ctx, ctypes.byref(n)))
        return n.value

    def extract_arrays(self):
        """(centres [n, dim], radii [n], phases [n]) in rsa_grid::extract_spheres order."""
        n = ctypes.c_uint64()
        check(lib().rsa_gpu_extract(self._ctx, None, None, None, 0, ctypes.byref(n)))
        cnt = n.value
        pos = np.zeros(cnt * self.dim)
        rad = np.zeros(cnt)
        ph = np.zeros(cnt, dtype=np.uint64)
        check(lib().rsa_gpu_extract(self._ctx, dptr(pos), dptr(rad), ph.ctypes.data_as(_lib._U64), cnt,
                                    ctypes.byref(n)))
        return pos.reshape(cnt, self.dim), rad, ph

    def extract_spheres(self):
        pos, rad, ph = self.extract_arrays()
        return [Sphere(p.tolist(), r, int(f)) for p, r, f in zip(pos, rad, ph)]

    def compute_total_volume_of_spheres(self):
        v = ctypes.c_double()
        check(lib().rsa_gpu_volume_of_spheres(self._ctx, ctypes.byref(v)))
        return v.value

    def add(self, centers, radii, phases=None):
        """rsa_domain::add (rsa_domain.hxx:84): insert validated spheres."""
        c = f64(centers).reshape(-1)
        r = f64(radii)
        ph = np.ascontiguousarray(phases if phases is not None else np.zeros(len(r)), dtype=np.uint64)
        check(lib().rsa_gpu_add_spheres(self._ctx, len(r), dptr(c), dptr(r), ph.ctypes.data_as(_lib._U64)))

    def profile(self, level=1):
        """Enable (level 1: refine kernels, 2: all) and reset the per-kernel event timing; 0 disables."""
        check(lib().rsa_gpu_profile_enable(self._ctx, int(level)))

    def kernel_profile(self):
        """{kernel: dict(launches, seconds, bytes)} accumulated since profile(True)."""
        out = {}
        k = 0
        while True:
            name = ctypes.c_char_p()
            n = ctypes.c_int64()
            sec = ctypes.c_double()
            by = ctypes.c_double()
            rc = lib().rsa_gpu_kernel_profile(self._ctx, k, ctypes.byref(name), ctypes.byref(n), ctypes.byref(sec),
                                              ctypes.byref(by))
            if rc == 2:
                break
            check(rc)
            out[name.value.decode()] = dict(launches=n.value, seconds=sec.value, bytes=by.value)
            k += 1
        return out

    def learn_mode(self, trace=False, use=False, measure=False):
        """The learned predictor on this domain's proceeds: trace (log voxel and draw samples
        for training), use (order the coverage walk and the draw queue by the trained model),
        measure (outcome and wave counters without use)."""
        check(lib().rsa_gpu_learn_set_mode(self._ctx, int(bool(trace)), int(bool(use)), int(bool(measure))))

    def learn_stats(self):
        st = _lib.LearnStats()
        check(lib().rsa_gpu_learn_get_stats(self._ctx, ctypes.byref(st)))
        return st.as_dict()

    def voxels(self):
        """Current voxel list (int64 lattice coordinates) and the common voxel length."""
        n = ctypes.c_uint64()
        ln = np.zeros(self.dim)
        check(lib().rsa_gpu_voxels(self._ctx, None, dptr(ln), 0, ctypes.byref(n)))
        co = np.zeros(n.value * self.dim, dtype=np.int64)
        check(lib().rsa_gpu_voxels(self._ctx, co.ctypes.data_as(_lib._I64), dptr(ln), n.value, ctypes.byref(n)))
        return co.reshape(-1, self.dim), ln


class RadiusGenerator:
    """sac_de_billes::RadiusGenerator<DIM> (radius_generator.hxx:89-149).

    ``RadiusGenerator([(r, N, phase), ...], exclusion)`` or
    ``RadiusGenerator([(r, phi, phase), ...], volume, exclusion)``.
    """

    def __init__(self, classes, volume_or_exclusion=0.0, exclusion=None):
        if len(classes) == 0:
            raise RuntimeError("Empty radius generator!")
        self.radius = [float(c[0]) for c in classes]
        self.phase = [int(c[2]) for c in classes]
        if exclusion is None:  # (radius, count, phase) form
            self.count = [int(c[1]) for c in classes]
            self.fraction = None
            self.volume = None
            self.exclusion = float(volume_or_exclusion)
        else:
            self.fraction = [float(c[1]) for c in classes]
            self.volume = float(volume_or_exclusion)
            self.exclusion = float(exclusion)
            self.count = None

    def counts(self, dim):
        """compute_desired_radius_nb_phase (radius_generator.ixx:71-86)."""
        if self.count is not None:
            return list(self.count)
        import math
        mult = 1.0
        if dim % 2 == 0:
            for k in range(dim // 2):
                mult *= (2 * math.pi) / (dim - 2 * k)
        else:
            mult = 1.0 / math.pi
            for k in range((dim - 1) // 2 + 1):
                mult *= (2 * math.pi) / (dim - 2 * k)
        out = []
        for r, phi in zip(self.radius, self.fraction):
            p = {2: r * r, 3: r * (r * r), 4: (r * r) * (r * r)}[dim]
            out.append(int(self.volume * phi / (mult * p)))
        return out

    def _arrays(self, dim):
        return (f64(self.radius), np.ascontiguousarray(self.counts(dim), dtype=np.uint64),
                np.ascontiguousarray(self.phase, dtype=np.int32))


class LawRadiusGenerator:
    """RadiusGenerator<DIM>(min_radius, max_radius, RandomRadiusGenerator*, n, exclusion)
    (radius_generator.hxx:62-70, radius_generator.ixx:38-48): `count` spheres whose
    radii are transform(u), u uniform in [0, 1) (radius_generator.hxx:12-38).
    `transform` maps a NumPy array of u to an array of radii in [min_radius, max_radius]
    and must be a pure function of u; it is called from a library thread, once per draw."""

    def __init__(self, min_radius, max_radius, transform, count, phase=0, exclusion=0.0):
        self.min_radius = float(min_radius)
        self.max_radius = float(max_radius)
        self.transform = transform
        self.count = int(count)
        self.phase = int(phase)
        self.exclusion = float(exclusion)

    def _callback(self):
        tf = self.transform

        def law(u, r, n, user):
            uu = np.ctypeslib.as_array(u, shape=(n,))
            out = np.ctypeslib.as_array(r, shape=(n,))
            out[:] = np.asarray(tf(uu), dtype=np.float64)
        return _lib.RADIUS_LAW(law)


class Algo:
    """algorithm::rsa_algo<DIM> (operator_algorithm.hxx:18-109), voxel method."""

    def __init__(self, domain, radius_generator, nbshots_singledraw=6000, n_draw=10):
        self.domain = domain
        if isinstance(radius_generator, LawRadiusGenerator):
            g = radius_generator
            self._law = g._callback()  # kept alive with the algorithm
            domain._law = self._law   # and with the domain, whose context calls it
            check(lib().rsa_gpu_algo_init_law(domain._ctx, g.min_radius, g.max_radius, self._law, None,
                                              g.phase, g.count, g.exclusion, int(nbshots_singledraw)))
        else:
            r, c, p = radius_generator._arrays(domain.dim)
            check(lib().rsa_gpu_algo_init(domain._ctx, len(r), dptr(r), c.ctypes.data_as(_lib._U64),
                                          p.ctypes.data_as(_lib._I), radius_generator.exclusion,
                                          int(nbshots_singledraw)))
        self.last_stats = None

    def proceed(self, seed, max_draws=-1):
        st = Stats()
        check(lib().rsa_gpu_proceed(self.domain._ctx, int(seed), int(max_draws), ctypes.byref(st)))
        self.last_stats = st.as_dict()
        return self.last_stats

    def reset_radius_generator(self, radius_generator):
        """rsa_algo::reset_radius_generator (operator_algorithm.ixx:238-250): classes or a custom law."""
        if isinstance(radius_generator, LawRadiusGenerator):
            g = radius_generator
            self._law = g._callback()
            self.domain._law = self._law
            check(lib().rsa_gpu_reset_law(self.domain._ctx, g.min_radius, g.max_radius, self._law, None, g.phase,
                                          g.count, g.exclusion))
            return
        r, c, p = radius_generator._arrays(self.domain.dim)
        check(lib().rsa_gpu_reset_classes(self.domain._ctx, len(r), dptr(r), c.ctypes.data_as(_lib._U64),
                                          p.ctypes.data_as(_lib._I), radius_generator.exclusion))

    def set_miss_rate(self, m):
        check(lib().rsa_gpu_set_miss_rate(self.domain._ctx, float(m)))


def uniform_generate(domain, rad_or_generator, size=6000, n_draw=1, seed=0):
    """algorithm::uniform_generate<DIM, 1> (operator_algorithm.ixx:275-292)."""
    if isinstance(rad_or_generator, RadiusGenerator):
        gen = rad_or_generator
    else:
        gen = RadiusGenerator([(float(rad_or_generator), 1.0, 0)], domain.get_total_volume(), 0.0)
    algo = Algo(domain, gen, size, n_draw)
    algo.proceed(seed)
    return algo


def throw_spheres(L, desired_radius_volumeFraction_phase, seed, write_paraview=False):
    """user_interface::throw_spheres (test.hxx:13-32): box [0, L]; the reference
    passes volume 1 to the radius generator (kept)."""
    dim = len(L)
    gen = RadiusGenerator(desired_radius_volumeFraction_phase, 1.0, 0.0)
    domain = Domain([0.0] * dim, [float(v) for v in L], max(gen.radius))
    uniform_generate(domain, gen, 6000, 10, seed)
    return domain.extract_spheres()


_BIN_MAGIC = b"RSAMPIB1"


def write_binary(path, pos, rad, phase):
    """Lossless dump of spheres (the layout of write_binary in the C++ drop-in): "RSAMPIB1",
    uint32 dim, uint32 0, uint64 n, then per sphere dim float64 coordinates, float64 radius,
    uint64 phase, little-endian."""
    pos = np.asarray(pos, dtype=np.float64)
    n, dim = pos.shape
    rec = np.zeros(n, dtype=[("c", "<f8", (dim,)), ("r", "<f8"), ("p", "<u8")])
    rec["c"] = pos
    rec["r"] = np.asarray(rad, dtype=np.float64)
    rec["p"] = np.asarray(phase, dtype=np.uint64)
    with open(path, "wb") as f:
        f.write(_BIN_MAGIC + np.array([dim, 0], dtype="<u4").tobytes() + np.array([n], dtype="<u8").tobytes())
        f.write(rec.tobytes())


def read_binary(path):
    """(pos, rad, phase) of a file write_binary (C++ or Python) wrote."""
    with open(path, "rb") as f:
        head = f.read(24)
        if len(head) != 24 or head[:8] != _BIN_MAGIC:
            raise ValueError(f"{path} is not a sphere dump")
        dim = int(np.frombuffer(head[8:12], dtype="<u4")[0])
        n = int(np.frombuffer(head[16:24], dtype="<u8")[0])
        rec = np.frombuffer(f.read(), dtype=[("c", "<f8", (dim,)), ("r", "<f8"), ("p", "<u8")], count=n)
    return rec["c"].copy(), rec["r"].copy(), rec["p"].copy()


def _sphere_cap(dim, sides, radii):
    """Upper bound on the spheres of the boxes: disjoint balls fill at most the
    whole volume (plus one per box for rounding)."""
    vball = {2: np.pi, 3: 4.0 * np.pi / 3.0, 4: np.pi ** 2 / 2.0}[int(dim)]
    return int(sum(float(s) ** dim / (vball * float(r) ** dim) + 1 for s, r in zip(sides, radii))) + 16


def batch_run(dim, sides, radii, seeds, nbshots=6000, workers=None, extract=False):
    """Independent realizations, realization k being uniform_generate<dim, 1> on
    the periodic box [0, sides[k]]^dim with radius radii[k] and seed seeds[k]
    (operator_algorithm.ixx:275-292), run as a device work queue with `workers` of
    them in flight on one stream (None: the count autotune_batch() stored for this
    dimension and median box-to-radius ratio, else 64).  Returns a dict: counts,
    volumes, seconds and, with extract, per realization the (pos, rad) arrays in
    extraction order."""
    sides = f64(sides)
    radii = f64(radii)
    if workers is None:
        workers = _tuning.lookup_batch(dim, sides, radii) or 64
    seeds = np.ascontiguousarray(seeds, dtype=np.uint64)
    n = len(sides)
    if len(radii) != n or len(seeds) != n:
        raise ValueError("sides, radii and seeds must have the same length")
    counts = np.zeros(n, dtype=np.uint64)
    volumes = np.zeros(n, dtype=np.float64)
    secs = np.zeros(1, dtype=np.float64)
    if extract:
        # upper bound on the spheres of a box: its volume over the volume of the
        # cube inscribed in one sphere
        cap = _sphere_cap(dim, sides, radii)
        pos = np.zeros(cap * dim, dtype=np.float64)
        rad = np.zeros(cap, dtype=np.float64)
        pp, rp = dptr(pos), dptr(rad)
    else:
        cap, pp, rp = 0, None, None
    check(lib().rsa_gpu_batch_run(dim, n, dptr(sides), dptr(radii), seeds.ctypes.data_as(_lib._U64), int(nbshots),
                                  int(workers), counts.ctypes.data_as(_lib._U64), dptr(volumes), pp, rp, cap,
                                  dptr(secs)))
    out = {"counts": counts, "volumes": volumes, "seconds": float(secs[0])}
    if extract:
        res, off = [], 0
        for c in counts.astype(np.int64):
            res.append((pos[off * dim:(off + c) * dim].reshape(c, dim), rad[off:off + c]))  # views
            off += c
        out["spheres"] = res
    return out


def batch_queue(dim, sides, radii, seeds, nbshots=6000, active=64, extract=False, volumes=False):
    """The same realizations as batch_run, driven as a device work queue: up to
    `active` of them in flight on one stream, every draw round and refinement step
    of all of them one launch (rsa_gpu_batch_queue).  Returns batch_run's dict
    (volumes only when asked for) plus "stats" (launch-group counts, host time by
    phase)."""
    sides = f64(sides)
    radii = f64(radii)
    seeds = np.ascontiguousarray(seeds, dtype=np.uint64)
    n = len(sides)
    if len(radii) != n or len(seeds) != n:
        raise ValueError("sides, radii and seeds must have the same length")
    counts = np.zeros(n, dtype=np.uint64)
    vols = np.zeros(n, dtype=np.float64)
    if extract:
        cap = _sphere_cap(dim, sides, radii)
        pos = np.zeros(cap * dim, dtype=np.float64)
        rad = np.zeros(cap, dtype=np.float64)
        pp, rp = dptr(pos), dptr(rad)
    else:
        cap, pp, rp = 0, None, None
    st = _lib.BatchStats()
    check(lib().rsa_gpu_batch_queue(dim, n, dptr(sides), dptr(radii), seeds.ctypes.data_as(_lib._U64), int(nbshots),
                                    int(active), counts.ctypes.data_as(_lib._U64), dptr(vols) if volumes else None,
                                    pp, rp, cap, ctypes.byref(st)))
    out = {"counts": counts, "volumes": vols, "seconds": st.seconds, "stats": st.as_dict()}
    if extract:
        res, off = [], 0
        for c in counts.astype(np.int64):
            res.append((pos[off * dim:(off + c) * dim].reshape(c, dim), rad[off:off + c]))  # views
            off += c
        out["spheres"] = res
    return out


def _make_dim_aliases():
    g = globals()
    for d in (2, 3, 4):
        def mk_domain(d=d):
            class _Dom(Domain):
                def __init__(self, global_inf, global_sup, rad):
                    if len(global_inf) != d:
                        raise ValueError("Incoherent dimensions!")
                    super().__init__(global_inf, global_sup, rad)
            _Dom.__name__ = f"rsa_domain_{d}D"
            return _Dom
        g[f"rsa_domain_{d}D"] = mk_domain()
        g[f"radius_generator_{d}D"] = RadiusGenerator
        g[f"rsa_algo_{d}D"] = Algo
        g[f"sphere{d}D"] = Sphere
        g[f"throw_spheres_{d}D"] = throw_spheres


_make_dim_aliases()


# ------------------------------------------------------------ the learned predictor (process-wide model)

def _fptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def learn_reset(seed=1):
    """Fresh MLP parameters and an empty trace store."""
    check(lib().rsa_gpu_learn_reset(int(seed)))


def learn_samples():
    n = ctypes.c_int64()
    check(lib().rsa_gpu_learn_samples(ctypes.byref(n)))
    return n.value


def learn_train(steps=400, batch=1024, lr=3e-3, seed=0):
    """Adam on the trace store with the GPU forward/backward kernels: (first loss, last loss, samples)."""
    a, b = ctypes.c_double(), ctypes.c_double()
    n = ctypes.c_int64()
    check(lib().rsa_gpu_learn_train(int(steps), int(batch), float(lr), int(seed), ctypes.byref(a), ctypes.byref(b),
                                    ctypes.byref(n)))
    return a.value, b.value, n.value


def learn_params(set=None):
    """The model's FP32 parameters (get), or install `set`."""
    if set is not None:
        p = np.ascontiguousarray(set, dtype=np.float32)
        assert p.size == _lib.NPAR
        check(lib().rsa_gpu_learn_params(None, _fptr(p)))
        return p
    out = np.zeros(_lib.NPAR, dtype=np.float32)
    check(lib().rsa_gpu_learn_params(_fptr(out), None))
    return out


def learn_gradcheck(x, t, params):
    """(gradients of the mean loss, loss) of one GPU forward/backward pass over the samples."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    t = np.ascontiguousarray(t, dtype=np.uint8)
    p = np.ascontiguousarray(params, dtype=np.float32)
    g = np.zeros(_lib.NPAR, dtype=np.float32)
    loss = ctypes.c_double()
    check(lib().rsa_gpu_learn_gradcheck(len(x), _fptr(x), t.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), _fptr(p),
                                        _fptr(g), ctypes.byref(loss)))
    return g, loss.value


def learn_forward(x, params):
    x = np.ascontiguousarray(x, dtype=np.float32)
    p = np.ascontiguousarray(params, dtype=np.float32)
    out = np.zeros((len(x), 2), dtype=np.float32)
    check(lib().rsa_gpu_learn_forward(len(x), _fptr(x), _fptr(p), _fptr(out)))
    return out
