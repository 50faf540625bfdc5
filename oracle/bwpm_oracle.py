"""FP64 CPU oracle of the BWPM bidirectional slab-stepping path.

TEST INFRASTRUCTURE ONLY. This module is the checker, never the product: only `tests/`,
`__graft_entry__.smoke()` and the `cpu_baseline` leg of `bench.py` may import it, and the
`bwpm_amd` package never does (its GPU path fails loudly when the HIP library is missing).

It restates, in vectorised NumPy complex128, the algorithm of the reference's CPU core:

- `fftfreq`                          cpp/src/solver.cpp:17-24
- `longitudinal_wavenumber`          cpp/src/interface.cpp:8-18   (src/bwpm/interface.py:26-48)
- `fresnel_coefficients`             cpp/src/interface.cpp:20-28  (src/bwpm/interface.py:68-90)
- `transmission_matrix`              cpp/src/interface.cpp:30-40  (src/bwpm/interface.py:93-123)
- `reflection_matrix`                cpp/src/interface.cpp:42-52  (src/bwpm/interface.py:126-151)
- `Solver.op` (table 1)              cpp/src/solver.cpp:73-106
- `Solver.apply`                     cpp/src/solver.cpp:384-417
- `store` / `add` (sparse sources)   cpp/src/solver.cpp:419-442
- `gmres`                            cpp/src/solver.cpp:130-207
- `Solver.born_sources`              cpp/src/solver.cpp:215-260
- `Solver.born_operator`             cpp/src/solver.cpp:262-311
- `Solver.born_step`                 cpp/src/solver.cpp:313-350
- `Solver.born_emission`             cpp/src/solver.cpp:352-382
- `Solver.run`                       cpp/src/solver.cpp:444-528
- `power_flux` (2D)                  src/bwpm/fpwpm.py:205-216, residual src/bwpm/fast.py:148-155
- `port_power` (3D)                  src/bwpm/adapter.py:267-284

The FFT is numpy.fft (unnormalised forward, 1/N inverse, as pocketfft in the reference core).

How each function is pinned (tests/test_oracle.py): the seven golden fixtures of the reference
(tests/golden/*.result.json, rel 1e-12), the reference's own CPU core compiled from its sources
into oracle/_ref (random volumes incl. silver, grazing points, δn, 3 iterations), and closed
forms (homogeneous plane wave, R + T = 1 for lossless media, the transfer-matrix solution of a
planar stack, the Airy reflectance of a slab, exact absorption e^{-2k0 κ L}, flux conservation).
"""

from __future__ import annotations

import numpy as np

GRAZING_TOLERANCE = 1e-12  # cpp/include/bwpm/interface.hpp:17
GMRES_TOLERANCE = 1e-12  # cpp/src/solver.cpp:209
GMRES_RESTART = 30  # cpp/src/solver.cpp:341
GMRES_MAX_RESTARTS = 100  # cpp/src/solver.cpp:341
GAUSS = ((1.0 - 1.0 / np.sqrt(3.0)) / 2.0, (1.0 + 1.0 / np.sqrt(3.0)) / 2.0)  # solver.cpp:212

FWD_T, FWD_R, BWD_T, BWD_R = 0, 1, 2, 3  # cpp/include/bwpm/solver.hpp:140


def fftfreq(n: int, d: float) -> np.ndarray:
    """2π·(m·(1/(n·d))) in numpy order; the rounding of 2π·numpy.fft.fftfreq (solver.cpp:17-24)."""
    i = np.arange(n)
    m = np.where(i < (n + 1) // 2, i, i - n).astype(np.float64)
    val = 1.0 / (float(n) * d)
    return 2.0 * np.pi * (m * val)


def longitudinal_wavenumber(index: complex, k0: float, k_perp_sq) -> np.ndarray:
    """k_z = sqrt(ñ²k₀² − k⊥²), Im ≥ 0 (Re ≥ 0 if Im = 0), grazing rule (interface.cpp:8-18)."""
    index = complex(index)
    n2k2 = index * index * (k0 * k0)
    q = np.asarray(n2k2 - np.asarray(k_perp_sq, dtype=np.float64), dtype=np.complex128)
    q = np.where(np.abs(q) <= GRAZING_TOLERANCE * abs(n2k2), 0.0 + 0.0j, q)
    kz = np.sqrt(q)
    flip = (kz.imag < 0.0) | ((kz.imag == 0.0) & (kz.real < 0.0))
    kz = np.where(flip, -kz, kz)
    tiny = 1e-12 * k0
    return np.where(np.abs(kz) < tiny, 1j * tiny, kz)


def fresnel_coefficients(kx, ky, n1: complex, n2: complex, k0: float):
    """(r_s, r_p, t_s, t_p, kz_in, kz_out), Convention A (interface.cpp:20-28)."""
    n1, n2 = complex(n1), complex(n2)
    kx = np.asarray(kx, dtype=np.float64)
    ky = np.asarray(ky, dtype=np.float64)
    kp2 = kx * kx + ky * ky
    kz1 = longitudinal_wavenumber(n1, k0, kp2)
    kz2 = longitudinal_wavenumber(n2, k0, kp2)
    q1, q2 = kz1 / k0, kz2 / k0
    den_p = n2 * n2 * q1 + n1 * n1 * q2
    return (
        (q1 - q2) / (q1 + q2),
        (n2 * n2 * q1 - n1 * n1 * q2) / den_p,
        2.0 * q1 / (q1 + q2),
        2.0 * n1 * n2 * q1 / den_p,
        kz1,
        kz2,
    )


def _stack22(m00, m01, m10, m11) -> np.ndarray:
    return np.stack([np.stack([m00, m01], -1), np.stack([m10, m11], -1)], -2)


def transmission_matrix(kx, ky, n1: complex, n2: complex, k0: float) -> np.ndarray:
    """T of wende2023 eqs. (2), (3), shape (..., 2, 2); k⊥ = 0 → t_s·I (interface.cpp:30-40)."""
    r_s, r_p, t_s, t_p, kz_in, kz_out = fresnel_coefficients(kx, ky, n1, n2, k0)
    kx = np.asarray(kx, dtype=np.float64)
    ky = np.asarray(ky, dtype=np.float64)
    kp2 = kx * kx + ky * ky
    kr, ks = complex(n1) * k0, complex(n2) * k0
    zero = kp2 == 0.0
    with np.errstate(divide="ignore", invalid="ignore"):
        gamma = kz_in * kz_out + kz_out * kp2 / kz_in
        scale = 1.0 / (kp2 * kr * ks)
        off = (t_p * gamma - t_s * kr * ks) * kx * ky * scale
        m00 = (t_p * kx * kx * gamma + t_s * ky * ky * kr * ks) * scale
        m11 = (t_p * ky * ky * gamma + t_s * kx * kx * kr * ks) * scale
    m00 = np.where(zero, t_s, m00)
    m11 = np.where(zero, t_s, m11)
    off = np.where(zero, 0.0, off)
    return _stack22(m00, off, off, m11)


def reflection_matrix(kx, ky, n1: complex, n2: complex, k0: float) -> np.ndarray:
    """R of wende2023 eqs. (8), (9), shape (..., 2, 2); k⊥ = 0 → r_s·I (interface.cpp:42-52)."""
    r_s, r_p, t_s, t_p, kz_in, kz_out = fresnel_coefficients(kx, ky, n1, n2, k0)
    kx = np.asarray(kx, dtype=np.float64)
    ky = np.asarray(ky, dtype=np.float64)
    kp2 = kx * kx + ky * ky
    kr = complex(n1) * k0
    zero = kp2 == 0.0
    with np.errstate(divide="ignore", invalid="ignore"):
        gamma = kz_in * kz_in + kp2
        scale = 1.0 / (kp2 * kr * kr)
        off = -(r_p * gamma + r_s * kr * kr) * kx * ky * scale
        m00 = (-r_p * kx * kx * gamma + r_s * ky * ky * kr * kr) * scale
        m11 = (-r_p * ky * ky * gamma + r_s * kx * kx * kr * kr) * scale
    m00 = np.where(zero, r_s, m00)
    m11 = np.where(zero, r_s, m11)
    off = np.where(zero, 0.0, off)
    return _stack22(m00, off, off, m11)


def gmres(apply, b: np.ndarray, x: np.ndarray, tol: float, restart: int, max_restarts: int):
    """Restarted GMRES, modified Gram–Schmidt, complex Givens (solver.cpp:130-207).

    Returns (converged, x, inner_iterations)."""
    x = x.copy()
    bnorm = np.sqrt(np.real(np.vdot(b, b)))
    if bnorm == 0.0:
        return True, np.zeros_like(x), 0
    m = restart
    total = 0
    for _ in range(max_restarts):
        w = apply(x)
        v0 = b - w
        beta = np.sqrt(np.real(np.vdot(v0, v0)))
        if beta <= tol * bnorm:
            return True, x, total
        basis = [v0 / beta]
        h = np.zeros((m + 1, m), dtype=np.complex128)
        g = np.zeros(m + 1, dtype=np.complex128)
        cs = np.zeros(m, dtype=np.complex128)
        sn = np.zeros(m, dtype=np.complex128)
        g[0] = beta
        used = 0
        done = False
        for j in range(m):
            w = apply(basis[j])
            total += 1
            for i in range(j + 1):
                hij = np.vdot(basis[i], w)
                h[i, j] = hij
                w = w - hij * basis[i]
            hnext = np.sqrt(np.real(np.vdot(w, w)))
            basis.append(w / hnext if hnext > 0.0 else np.zeros_like(w))
            for i in range(j):
                a, c = h[i, j], h[i + 1, j]
                h[i, j] = cs[i] * a + sn[i] * c
                h[i + 1, j] = -np.conj(sn[i]) * a + cs[i] * c
            a = h[j, j]
            r = np.sqrt(abs(a) ** 2 + hnext * hnext)
            if abs(a) == 0.0:
                cs[j], sn[j] = 0.0, 1.0
            else:
                cs[j] = abs(a) / r
                sn[j] = (a / abs(a)) * hnext / r
            h[j, j] = cs[j] * a + sn[j] * hnext
            g[j + 1] = -np.conj(sn[j]) * g[j]
            g[j] = cs[j] * g[j]
            used = j + 1
            if abs(g[j + 1]) <= tol * bnorm or hnext == 0.0:
                done = True
                break
        y = np.zeros(used, dtype=np.complex128)
        for i in range(used - 1, -1, -1):
            s = g[i]
            for k in range(i + 1, used):
                s -= h[i, k] * y[k]
            y[i] = s / h[i, i]
        for i in range(used):
            x = x + y[i] * basis[i]
        if done:
            w = apply(x)
            res = np.sqrt(np.sum(np.abs(b - w) ** 2))
            if res <= tol * bnorm * 10.0:
                return True, x, total
    return False, x, total


class Solver:
    """FP64 restatement of bwpm::Solver (cpp/include/bwpm/solver.hpp:127-178).

    Arguments follow the reference binding `_core.Solver(ny, nx, dy, dx, dz, k0, indices, labels,
    index_before, mask, threads, fluctuation)` (cpp/python/bindings.cpp:139-151); `threads` is
    accepted and ignored. Fields are (ny, nx, 2) complex128 with components (E_x, E_y).
    """

    def __init__(self, ny, nx, dy, dx, dz, k0, indices, labels, index_before, mask=None,
                 threads=1, fluctuation=None, gmres_max_restarts=GMRES_MAX_RESTARTS):
        self.ny, self.nx = int(ny), int(nx)
        self.dy, self.dx, self.dz, self.k0 = float(dy), float(dx), float(dz), float(k0)
        self.indices = np.asarray(indices, dtype=np.complex128).ravel()
        labels = np.asarray(labels)
        if labels.ndim != 3 or labels.shape[1:] != (self.ny, self.nx):
            raise ValueError("labels must have shape (nz, ny, nx)")
        self.labels = labels.astype(np.int64)
        self.nz = labels.shape[0]
        if self.labels.size and (self.labels.min() < 0 or self.labels.max() >= self.indices.size):
            raise ValueError("labels must index into `indices`")
        self.index_before = complex(index_before)
        self.mask = None if mask is None else np.asarray(mask, dtype=np.float64)
        if self.mask is not None and self.mask.size != self.ny * self.nx:
            raise ValueError("mask must have shape (ny, nx)")
        if self.mask is not None:
            self.mask = self.mask.reshape(self.ny, self.nx)
        kx = fftfreq(self.nx, self.dx)
        ky = fftfreq(self.ny, self.dy)
        self.kx = np.broadcast_to(kx[None, :], (self.ny, self.nx)).copy()
        self.ky = np.broadcast_to(ky[:, None], (self.ny, self.nx)).copy()
        self.gmres_max_restarts = gmres_max_restarts
        self.gmres_iterations: list[int] = []
        self.fluct = None
        self.host: list[complex | None] = [None] * self.nz
        if fluctuation is not None:
            f = np.asarray(fluctuation, dtype=np.complex128)
            if f.shape != self.labels.shape:
                raise ValueError("fluctuation must have the shape of labels (nz, ny, nx)")
            self.fluct = f
            for j in range(self.nz):
                where = f[j] != 0
                found = {complex(v) for v in self.indices[self.labels[j][where]]}
                if len(found) > 1:
                    raise ValueError(f"slab {j}: fluctuation must lie in one host index")
                self.host[j] = found.pop() if found else None
        self._cache: dict = {}

    # -- planes and operators ---------------------------------------------------------------
    def plane_labels(self, p: int) -> np.ndarray:
        """Plane 0: label −1 (index_before); plane p: labels of slab p − 1 (solver.cpp:65-67)."""
        if p == 0:
            return np.full((self.ny, self.nx), -1, dtype=np.int64)
        return self.labels[p - 1]

    def index_of(self, label: int) -> complex:
        return self.index_before if label < 0 else complex(self.indices[label])

    def plane_indices(self, p: int) -> np.ndarray:
        lab = self.plane_labels(p)
        if p == 0:
            return np.full(lab.shape, self.index_before, dtype=np.complex128)
        return self.indices[lab]

    def op(self, kind: int, r: int, s: int) -> np.ndarray:
        """Table 1 of wende2023, (ny, nx, 2, 2) (solver.cpp:73-106)."""
        key = (kind, r, s)
        if key in self._cache:
            return self._cache[key]
        nr, ns = self.index_of(r), self.index_of(s)
        kp2 = self.kx * self.kx + self.ky * self.ky
        if kind == FWD_T:
            a = transmission_matrix(self.kx, self.ky, nr, ns, self.k0)
            phase = np.exp(1j * longitudinal_wavenumber(ns, self.k0, kp2) * self.dz)
        elif kind == FWD_R:
            a = reflection_matrix(self.kx, self.ky, nr, ns, self.k0)
            phase = np.ones(kp2.shape, dtype=np.complex128)
        elif kind == BWD_T:
            a = transmission_matrix(self.kx, self.ky, nr, ns, self.k0)
            phase = np.exp(1j * longitudinal_wavenumber(nr, self.k0, kp2) * self.dz)
        else:
            a = reflection_matrix(self.kx, self.ky, nr, ns, self.k0)
            phase = np.exp(2.0 * 1j * longitudinal_wavenumber(nr, self.k0, kp2) * self.dz)
        m = a * phase[..., None, None]
        self._cache[key] = m
        return m

    @staticmethod
    def fft(data: np.ndarray, forward: bool) -> np.ndarray:
        """2D c2c over (ny, nx) per component; inverse scaled by 1/N (solver.cpp:108-125)."""
        if forward:
            return np.fft.fft2(data, axes=(0, 1))
        return np.fft.ifft2(data, axes=(0, 1))

    def apply(self, kind: int, e: np.ndarray, p_r: int, p_s: int) -> np.ndarray:
        """Σ_pairs I_r,s ⊙ F⁻¹{M F{e}}; reflections skip equal indices; mask on transmissions
        (solver.cpp:384-417)."""
        lr, ls = self.plane_labels(p_r), self.plane_labels(p_s)
        pairs = sorted(set(zip(lr.ravel().tolist(), ls.ravel().tolist())))
        spectrum = self.fft(e, True)
        out = np.zeros_like(e)
        reflection = kind in (FWD_R, BWD_R)
        for r, s in pairs:
            if reflection and self.index_of(r) == self.index_of(s):
                continue
            m = self.op(kind, r, s)
            work = np.einsum("...ij,...j->...i", m, spectrum)
            work = self.fft(work, False)
            sel = (lr == r) & (ls == s)
            out[sel] = work[sel]
        if not reflection and self.mask is not None:
            out = out * self.mask[..., None]
        return out

    # -- Born machinery ----------------------------------------------------------------------
    def _born_k(self, slab: int):
        host = self.host[slab]
        host2k2 = host * host * self.k0 * self.k0
        kabs2 = abs(host2k2)
        kp2 = self.kx * self.kx + self.ky * self.ky
        kz = longitudinal_wavenumber(host, self.k0, kp2)
        prop = (kp2 < kabs2) & (np.abs(host2k2 - kp2) > GRAZING_TOLERANCE * kabs2)
        return host, kz, prop

    def born_sources(self, e: np.ndarray, direction: int, slab: int) -> np.ndarray:
        """Ŝ_m = F{V·E_m} at the two Gauss points, (2, ny, nx, 3) (solver.cpp:215-260)."""
        host, kz, prop = self._born_k(slab)
        spectrum = self.fft(e, True)
        ex, ey = spectrum[..., 0], spectrum[..., 1]
        safe = np.where(prop, kz, 1.0)
        ez = np.where(prop, -float(direction) * (self.kx * ex + self.ky * ey) / safe, 0.0)
        full = np.stack([ex, ey, ez], axis=-1)
        scale = 2.0 * host * self.k0 * self.k0
        v = scale * self.fluct[slab]
        out = np.empty((2, self.ny, self.nx, 3), dtype=np.complex128)
        for m in range(2):
            distance = (GAUSS[m] if direction > 0 else 1.0 - GAUSS[m]) * self.dz
            phase = np.where(prop, np.exp(1j * kz * distance), 0.0)
            work = self.fft(full * phase[..., None], False)
            work = v[..., None] * work
            out[m] = self.fft(work, True)
        return out

    def born_operator(self, spectrum: np.ndarray, direction: int, slab: int) -> np.ndarray:
        """G_d·S = i dz/(2k_z)·[P(k_d)·F{V·e}]⊥ (solver.cpp:262-311)."""
        host, kz, prop = self._born_k(slab)
        k2 = (host * self.k0) ** 2
        d = float(direction)
        safe = np.where(prop, kz, 1.0)
        ex = np.where(prop, spectrum[..., 0], 0.0)
        ey = np.where(prop, spectrum[..., 1], 0.0)
        ez = np.where(prop, -d * (self.kx * ex + self.ky * ey) / safe, 0.0)
        work = self.fft(np.stack([ex, ey, ez], axis=-1), False)
        scale = 2.0 * host * self.k0 * self.k0
        work = (scale * self.fluct[slab])[..., None] * work
        work = self.fft(work, True)
        green = np.where(prop, 1j * self.dz / (2.0 * safe), 0.0)
        s0, s1, s2 = work[..., 0], work[..., 1], work[..., 2]
        longitudinal = (self.kx * s0 + self.ky * s1 + d * kz * s2) / k2
        out = np.stack([green * (s0 - self.kx * longitudinal), green * (s1 - self.ky * longitudinal)],
                       axis=-1)
        return np.where(prop[..., None], out, 0.0)

    def born_step(self, e: np.ndarray, direction: int, slab: int) -> np.ndarray:
        """Strang–Cayley increment P_½(1 − G/2)⁻¹(1 + G/2)P_½ − 1, real space (solver.cpp:313-350)."""
        host, kz, prop = self._born_k(slab)
        half = np.where(prop, np.exp(1j * kz * (self.dz / 2.0)), 0.0)
        x = self.fft(e, True) * half[..., None]
        shape = x.shape
        rhs = self.born_operator(x, direction, slab).ravel()

        def apply(v):
            return v - 0.5 * self.born_operator(v.reshape(shape), direction, slab).ravel()

        ok, y, its = gmres(apply, rhs, rhs.copy(), GMRES_TOLERANCE, GMRES_RESTART,
                           self.gmres_max_restarts)
        self.gmres_iterations.append(its)
        if not ok:
            raise RuntimeError("GMRES did not converge in the Born step")
        y = y.reshape(shape)
        if direction > 0:
            f = half
        else:
            with np.errstate(divide="ignore", invalid="ignore"):
                f = np.where(half == 0.0, half, 1.0 / np.where(half == 0.0, 1.0, half))
        return self.fft(y * f[..., None], False)

    def born_emission(self, sources: np.ndarray, direction: int, slab: int) -> np.ndarray:
        """ΔE⊥ = F⁻¹{Σ_m ½e^{i d′k_z(dz−ζ_m)}·i dz/(2k_z)·[P(k_d′)Ŝ_m]⊥} (solver.cpp:352-382)."""
        host, kz, prop = self._born_k(slab)
        k2 = (host * self.k0) ** 2
        d = float(direction)
        safe = np.where(prop, kz, 1.0)
        signed = d * kz
        total = np.zeros((self.ny, self.nx, 2), dtype=np.complex128)
        for m in range(2):
            green = 0.5 * np.exp(1j * signed * ((1.0 - GAUSS[m]) * self.dz)) * (
                1j * self.dz / (2.0 * safe))
            s = sources[m]
            longitudinal = (self.kx * s[..., 0] + self.ky * s[..., 1] + signed * s[..., 2]) / k2
            total[..., 0] += np.where(prop, green * (s[..., 0] - self.kx * longitudinal), 0.0)
            total[..., 1] += np.where(prop, green * (s[..., 1] - self.ky * longitudinal), 0.0)
        return self.fft(total, False)

    # -- bidirectional driver ----------------------------------------------------------------
    def interfaces(self) -> list[np.ndarray]:
        """Pixels where the index changes between plane p and p + 1 (solver.cpp:449-456)."""
        return [self.plane_indices(p) != self.plane_indices(p + 1) for p in range(self.nz)]

    def run(self, field, iterations: int, keep=()) -> dict:
        """Bidirectional iterations; `iterations = 0` is the forward FPWPM (solver.cpp:444-528)."""
        n = self.ny * self.nx
        field = np.asarray(field, dtype=np.complex128)
        if field.size != 2 * n:
            raise ValueError("field must have shape (ny, nx, 2)")
        field = field.reshape(self.ny, self.nx, 2)
        keep = set(int(k) for k in keep)
        iface = self.interfaces()
        anyp = [bool(m.any()) for m in iface]
        nz = self.nz
        res = {"transmitted": None, "reflected": None, "transmitted_iter": [],
               "reflected_iter": [], "forward": {}, "backward": {}, "stored_values": 0}
        src_fwd: dict[int, tuple] = {}
        born_fwd: dict[int, np.ndarray] = {}

        def store(v, where):
            idx = np.flatnonzero(where.ravel())
            return idx, v.reshape(n, 2)[idx].copy()

        def add(v, s):
            flat = v.reshape(n, 2)
            flat[s[0]] += s[1]

        for _ in range(max(iterations, 1)):
            src_bwd: dict[int, tuple] = {}
            born_bwd: dict[int, np.ndarray] = {}
            e = field.copy()
            if 0 in src_fwd:
                add(e, src_fwd[0])
            if 0 in keep:
                res["forward"][0] = e.copy()
            for p in range(nz):
                if iterations > 0 and anyp[p]:
                    src_bwd[p] = store(self.apply(FWD_R, e, p, p + 1), iface[p])
                nxt = self.apply(FWD_T, e, p, p + 1)
                if self.fluct is not None and self.host[p] is not None:
                    inc = self.born_step(e, +1, p)
                    m = 1.0 if self.mask is None else self.mask[..., None]
                    nxt = nxt + inc * m
                    if iterations > 0:
                        born_bwd[p + 1] = self.born_emission(self.born_sources(e, +1, p), -1, p)
                e = nxt
                if p + 1 in src_fwd:
                    add(e, src_fwd[p + 1])
                if p + 1 in born_fwd:
                    e = e + born_fwd[p + 1]
                if p + 1 in keep:
                    res["forward"][p + 1] = e.copy()
            res["transmitted"] = e
            if iterations == 0:
                break
            src_fwd = {}
            born_fwd = {}
            e = np.zeros((self.ny, self.nx, 2), dtype=np.complex128)
            if nz in src_bwd:
                add(e, src_bwd[nz])
            if nz in born_bwd:
                e = e + born_bwd[nz]
            if nz in keep:
                res["backward"][nz] = e.copy()
            for p in range(nz, 0, -1):
                if self.fluct is not None and self.host[p - 1] is not None:
                    born_fwd[p] = self.born_emission(self.born_sources(e, -1, p - 1), +1, p - 1)
                    e = e + self.born_step(e, -1, p - 1)
                if anyp[p - 1]:
                    src_fwd[p] = store(self.apply(BWD_R, e, p, p - 1), iface[p - 1])
                e = self.apply(BWD_T, e, p, p - 1)
                if p - 1 in src_bwd:
                    add(e, src_bwd[p - 1])
                if p - 1 in born_bwd:
                    e = e + born_bwd[p - 1]
                if p - 1 in keep:
                    res["backward"][p - 1] = e.copy()
            res["reflected"] = e
            res["transmitted_iter"].append(res["transmitted"])
            res["reflected_iter"].append(res["reflected"])
            stored = sum(s[0].size for s in src_bwd.values()) + sum(
                s[0].size for s in src_fwd.values())
            res["stored_values"] = max(res["stored_values"], stored)
        if res["reflected"] is None:
            res["reflected"] = np.zeros((self.ny, self.nx, 2), dtype=np.complex128)
        return res


def power_flux(field_y2, ny: int, dy: float, k0: float, index: complex, direction: int = 1) -> float:
    """2D Poynting flux Σ Re(k_z)(|Ê⊥|² + |Ê_z|²)·ny·dy (src/bwpm/fpwpm.py:205-216)."""
    ky = fftfreq(ny, dy)
    spectrum = np.fft.fft(np.asarray(field_y2), axis=0) / ny
    kz = longitudinal_wavenumber(index, k0, ky ** 2)
    with np.errstate(divide="ignore", invalid="ignore"):
        ez = np.where(kz != 0, -direction * ky * spectrum[:, 1] / kz, 0.0)
    magnitude = np.sum(np.abs(spectrum) ** 2, axis=1) + np.abs(ez) ** 2
    return float(np.sum(np.real(kz) * magnitude) * ny * dy)


def energy_residual_2d(field_in, transmitted_iter, reflected_iter, ny, dy, k0, index_before,
                       index_last) -> list[float]:
    """ε = 1 − (P_R + P_T)/P_in per iteration (src/bwpm/fast.py:148-155)."""
    p_in = power_flux(field_in, ny, dy, k0, index_before)
    out = []
    for t, r in zip(transmitted_iter, reflected_iter):
        p_r = power_flux(r, ny, dy, k0, index_before, direction=-1)
        p_t = power_flux(t, ny, dy, k0, index_last)
        out.append(1.0 - (p_r + p_t) / p_in if p_in > 0 else float("nan"))
    return out


def port_power(field, nx, ny, dx, dy, k0, index, direction=1) -> float:
    """3D Poynting flux through a port (src/bwpm/adapter.py:267-284)."""
    kx = np.broadcast_to(fftfreq(nx, dx)[None, :], (ny, nx))
    ky = np.broadcast_to(fftfreq(ny, dy)[:, None], (ny, nx))
    spectrum = np.fft.fft2(np.asarray(field), axes=(0, 1))
    kz = longitudinal_wavenumber(index, k0, kx * kx + ky * ky)
    with np.errstate(divide="ignore", invalid="ignore"):
        ez = np.where(kz != 0, -direction * (kx * spectrum[..., 0] + ky * spectrum[..., 1]) / kz,
                      0.0)
    magnitude = np.sum(np.abs(spectrum) ** 2, axis=-1) + np.abs(ez) ** 2
    return float(np.sum(np.real(kz) * magnitude) * dx * dy / (nx * ny))
