"""Seeded synthetic MRI problems (trajectories, coils, phantoms, field terms).

These produce the inputs of the benchmark workloads and of the tests; they
are plain NumPy (FP64) and deterministic for a given seed (PCG64).

* :func:`spiral`         Archimedean variable-density spiral
  ``k(tau) = 0.5 tau^1.5 exp(i 2 pi n_turns tau)``, cycles/pixel.
* :func:`gaussian_coils` Gaussian sensitivities (sigma = 0.4 FOV) centred on a
  ring of radius 0.6 FOV, random constant phase per coil.
* :func:`ellipse_phantom` sum of random ellipses, amplitudes U(0.1, 1).
* :func:`nufft_problem`  2-D non-Cartesian SENSE problem (image 128^2, 8 coils,
  16384-sample spiral, data at SNR 40).
* :func:`highorder_problem` 2-D expanded-encoding problem (220^2, 32 coils,
  24000 samples at 1 us, 16 solid harmonics up to 3rd order on a 0.2 m FOV
  with a smooth drift, a smooth +-100 Hz B0 map; optional 70 kHz zeroth-order
  term that drives |Phi| beyond 1e4 rad).
* :func:`batch_problems` uneven multi-slice batch (N in 96..160 even,
  M in 6000..20000, 16 coils).
"""
import numpy as np

HARM_ORDER = np.array([0] + [1] * 3 + [2] * 5 + [3] * 7 + [4] * 9 + [5] * 11)


def grid_size(n, os=1.5):
    """Oversampled grid side: 2*ceil(n*os/2), raised by 2 until its largest
    prime factor is at most 7 (the rule of nufftOp.m:146-152)."""
    g = int(np.ceil(n * os / 2.0)) * 2

    def maxp(v):
        best, p = 1, 2
        while p * p <= v:
            while v % p == 0:
                best, v = p, v // p
            p += 1
        return max(best, v) if v > 1 else best
    while maxp(g) > 7:
        g += 2
    return g


def spiral(M, n_turns=64, dens_exp=1.5, kmax=0.5, wrap=True):
    """[M x 2] cycles/pixel; sample m at tau = m/(M-1).  With `wrap` a
    coordinate at +0.5 is moved to -0.5 (the same grid point)."""
    tau = np.arange(M) / (M - 1)
    k = kmax * tau ** dens_exp * np.exp(2j * np.pi * n_turns * tau)
    kl = np.stack([k.real, k.imag], axis=1)
    if wrap:
        kl = np.where(kl >= 0.5, kl - 1.0, kl)
    return kl


def spiral_dcf(M, n_turns=64, dens_exp=1.5, kmax=0.5):
    """Analytic density compensation |k| |dk/dtau| of :func:`spiral`,
    normalised to mean 1."""
    tau = np.arange(M) / (M - 1)
    r = kmax * tau ** dens_exp
    dr = kmax * dens_exp * tau ** (dens_exp - 1)
    dth = 2 * np.pi * n_turns
    speed = np.sqrt(dr ** 2 + (r * dth) ** 2)
    w = r * speed
    w[0] = w[1] * 0.5
    return w / w.mean()


def image_coords(N1, N2):
    """FOV-unit coordinates (i - N/2)/N of a column-major [N1, N2] image."""
    x = (np.arange(N1) - N1 / 2) / N1
    y = (np.arange(N2) - N2 / 2) / N2
    return np.meshgrid(x, y, indexing="ij")


def gaussian_coils(N1, N2, C, rng, sigma=0.4, radius=0.6):
    """[N1, N2, C] complex sensitivities."""
    X, Y = image_coords(N1, N2)
    phase = rng.uniform(0, 2 * np.pi, C)
    maps = np.empty((N1, N2, C), dtype=np.complex128)
    for c in range(C):
        cx, cy = radius * np.cos(2 * np.pi * c / C), radius * np.sin(2 * np.pi * c / C)
        maps[..., c] = np.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / (2 * sigma ** 2)) * np.exp(1j * phase[c])
    return maps


def ellipse_phantom(N1, N2, rng, n=10):
    X, Y = image_coords(N1, N2)
    im = np.zeros((N1, N2))
    for _ in range(n):
        cx, cy = rng.uniform(-0.3, 0.3, 2)
        a, b = rng.uniform(0.05, 0.3, 2)
        th = rng.uniform(0, np.pi)
        amp = rng.uniform(0.1, 1.0)
        xr = (X - cx) * np.cos(th) + (Y - cy) * np.sin(th)
        yr = -(X - cx) * np.sin(th) + (Y - cy) * np.cos(th)
        im += amp * ((xr / a) ** 2 + (yr / b) ** 2 <= 1)
    return im.astype(np.complex128)


def add_noise(y, rng, snr=40.0):
    sig = np.sqrt(np.mean(np.abs(y) ** 2)) / snr
    return y + sig / np.sqrt(2) * (rng.standard_normal(y.shape) + 1j * rng.standard_normal(y.shape))


def nufft_problem(seed=1, N=128, M=16384, C=8, n_turns=64, snr=40.0, with_data=True):
    """Inputs of the 2-D spiral CG-SENSE problem.  Returns a dict with kloc,
    maps, phantom and (optionally) noisy data computed with the FP64
    forward operator `forward` (a callable image -> [M, C]) if given."""
    rng = np.random.default_rng(seed)
    kloc = spiral(M, n_turns)
    maps = gaussian_coils(N, N, C, rng)
    ph = ellipse_phantom(N, N, rng)
    return {"N": N, "M": M, "C": C, "kloc": kloc, "maps": maps, "phantom": ph, "rng": rng,
            "dcf": spiral_dcf(M, n_turns)}


def highorder_problem(seed=2, N=220, K=24000, C=32, L=16, fov=0.2, z=0.02, dwell=1e-6,
                      n_turns=64, b0_hz=100.0, f0_hz=0.0):
    """Inputs of the 2-D expanded-encoding problem (see module docstring)."""
    rng = np.random.default_rng(seed)
    dx = fov / N
    gx = (np.arange(N) - N // 2) * dx
    X, Y = np.meshgrid(gx, gx, indexing="ij")
    grid = {"x": X, "y": Y, "z": np.full_like(X, z)}
    t = np.arange(K) * dwell
    k = spiral(K, n_turns, wrap=False)  # cycles/pixel
    kc = (k[:, 0] + 1j * k[:, 1]) * 2 * np.pi * N / fov  # rad/m
    h = np.zeros((L, K))
    h[1] = kc.real
    h[2] = kc.imag
    hmax = np.max(np.abs(h[1]))
    for l in range(L):
        d = np.zeros(K)
        for _ in range(5):
            f = rng.uniform(5, 200)
            ps = rng.uniform(0, 2 * np.pi)
            d += rng.uniform(0.2, 1.0) * np.sin(2 * np.pi * f * t + ps)
        target = 0.05 * hmax * (fov / 2) ** (1 - HARM_ORDER[l])
        h[l] += d / np.max(np.abs(d)) * target
    if f0_hz:
        h[0] += 2 * np.pi * f0_hz * t
    # smooth B0 map, +-b0_hz
    from scipy.ndimage import gaussian_filter
    b = gaussian_filter(rng.standard_normal((N, N)), 20, mode="wrap")
    b = b / np.max(np.abs(b)) * b0_hz
    b0 = 2 * np.pi * b
    maps = gaussian_coils(N, N, C, rng)
    ph = ellipse_phantom(N, N, rng)
    return {"N": N, "K": K, "C": C, "grid": grid, "t": t, "h": h, "b0": b0, "maps": maps,
            "phantom": ph, "rng": rng}


def nufft3_problem(seed=6, N=(128, 128, 64), C=8, per_partition=4096, n_turns=16):
    """Inputs of the 3-D stack-of-spirals SENSE problem: N3 partitions (kz on
    the Cartesian grid, cycles/pixel), each a 2-D variable-density spiral of
    `per_partition` samples (the spiral above, `n_turns` turns: about four
    times undersampled in-plane at 128 x 128) rotated by the golden angle from
    one partition to the next; C coils (the 2-D Gaussian-ring maps times a
    Gaussian z profile centred alternately in the two halves of the volume);
    a phantom of ellipses extruded over the middle 80 % of the partitions."""
    rng = np.random.default_rng(seed)
    N1, N2, N3 = N
    k2 = spiral(per_partition, n_turns, wrap=False)
    ga = np.pi * (3.0 - np.sqrt(5.0))
    kl = []
    for p in range(N3):
        c, s_ = np.cos(ga * p), np.sin(ga * p)
        kx = c * k2[:, 0] - s_ * k2[:, 1]
        ky = s_ * k2[:, 0] + c * k2[:, 1]
        kl.append(np.stack([kx, ky, np.full(per_partition, (p - N3 // 2) / N3)], axis=1))
    kloc = np.concatenate(kl)
    kloc = np.where(kloc >= 0.5, kloc - 1.0, kloc)
    z = (np.arange(N3) - N3 // 2) / N3
    maps2 = gaussian_coils(N1, N2, C, rng)
    prof = np.stack([np.exp(-(z - (0.25 if c % 2 else -0.25)) ** 2 / (2 * 0.4 ** 2)) for c in range(C)], axis=-1)
    maps = maps2[:, :, None, :] * prof[None, None, :, :]
    ph = ellipse_phantom(N1, N2, rng)[:, :, None] * (np.abs(z) < 0.4).astype(float)[None, None, :]
    return {"N": tuple(N), "C": C, "M": kloc.shape[0], "kloc": kloc, "maps": maps, "phantom": ph}


def delay_demo_problem(record, N=160, C=16, seed=5, delay=2.0):
    """Inputs of the trajectory-delay demo (demos/demo_autoDel.m:12-59) on
    synthetic images: a field-probe record (dict with phs_spha [K_in x 16],
    phs_conc [K_in x 4], datatime and tdwelltraj, as the reference's
    data_traj_spiral_R4.mat holds; acquisition from 50 us as the demo's
    spiral case), an N x N grid at 1.35 mm (the record's resolution), C
    Gaussian coils, an ellipse phantom and a smooth +-50 Hz B0 map in rad/us;
    the data are to be simulated with the coefficients delayed by `delay`."""
    rng = np.random.default_rng(seed)
    fov = 0.00135 * N
    gx = (np.arange(N) - N // 2) * (fov / N)
    X, Y = np.meshgrid(gx, gx, indexing="ij")
    grid = {"x": X, "y": Y, "z": np.full_like(X, 0.02)}
    from scipy.ndimage import gaussian_filter
    b = gaussian_filter(rng.standard_normal((N, N)), 15, mode="wrap")
    b0 = 2 * np.pi * 50.0 * b / np.max(np.abs(b)) * 1e-6
    maps = gaussian_coils(N, N, C, rng)
    ph = ellipse_phantom(N, N, rng)
    return {"N": N, "C": C, "grid": grid, "b0": b0, "maps": maps, "phantom": ph,
            "spha": np.ascontiguousarray(np.asarray(record["phs_spha"], np.float64)),
            "conc": np.ascontiguousarray(np.asarray(record["phs_conc"], np.float64)),
            "tdwell": float(np.asarray(record["tdwelltraj"]).reshape(-1)[0]),
            "datatime": np.asarray(record["datatime"], np.float64).reshape(-1) + 50.0, "delay": float(delay)}


def batch_problems(seed=4, nslice=64, C=16, Nrange=(96, 160), Mrange=(6000, 20000)):
    """Per-slice inputs of the uneven batch: list of dicts (N, M, kloc, maps, phantom)."""
    out = []
    for s in range(nslice):
        rng = np.random.default_rng(seed * 1000 + s)
        N = int(rng.integers(Nrange[0] // 2, Nrange[1] // 2 + 1)) * 2
        M = int(rng.integers(Mrange[0], Mrange[1] + 1))
        kloc = spiral(M, N // 2)
        maps = gaussian_coils(N, N, C, rng)
        ph = ellipse_phantom(N, N, rng)
        out.append({"N": N, "M": M, "kloc": kloc, "maps": maps, "phantom": ph, "rng": rng})
    return out
