"""Synthetic transport problems of the workload shapes (grids, frequency
mesh, source spectrum, opacity / re-emission tables).

Geometry follows the reference exactly so the walk is the reference's walk:
  axes            fillGrid, grid_mod.f90:530-549 (symmetric) and 567-596
  active map      setMotherGrid, grid_mod.f90:1225-1294 (z fastest numbering)
  geoCorr         fillGrid, grid_mod.f90:807-812
  source cell     setStarPosition, grid_mod.f90:3569-3603
  frequency mesh  initCartesianGrid, grid_mod.f90:213-266 (log mesh + edges)
  source CDF      setProbDen of a blackbody, continuum_mod.f90:418-474

The tables themselves are a documented physical stand-in for the reference's
ionizationDriver/emissionDriver (which are upstream of the transport path):
hydrogenic photoionisation cross sections of H0, He0, He+ and of the
dominant stages of C, N, O, Ne, S, a radially integrated on-the-spot
ionisation structure for a central source, Milne recombination continua plus
free-free emission for the re-emission CDFs, and a fixed line fraction.
"""
import numpy as np

f32 = np.float32
HC_K_RYD = 157893.94            # h c Ryd / k [K]
SERIES_EDGES = [0.0069, 0.0083, 0.01, 0.0123, 0.0156, 0.0204, 0.0278, 0.04, 0.0625,
                0.11117, 0.11610, 0.12248, 0.13732, 0.24763, 0.24994, 0.26630, 0.29189]
# (threshold [Ryd], sigma at threshold [cm^2], power-law index, element Z, stage)
EDGES = [
    (0.999434, 6.30e-18, 3.0, 1, 1), (1.80804, 7.40e-18, 2.0, 2, 1), (4.0, 1.58e-18, 3.0, 2, 2),
    (0.8277, 1.2e-17, 2.0, 6, 1), (1.7920, 4.6e-18, 2.5, 6, 2), (3.5250, 1.8e-18, 2.5, 6, 3),
    (1.0679, 1.1e-17, 2.0, 7, 1), (2.1753, 6.5e-18, 2.5, 7, 2), (3.4958, 2.7e-18, 2.5, 7, 3),
    (1.0005, 5.2e-18, 2.0, 8, 1), (2.5820, 7.3e-18, 2.5, 8, 2), (4.0343, 3.6e-18, 2.5, 8, 3),
    (1.5850, 4.9e-18, 2.0, 10, 1), (3.0200, 5.1e-18, 2.5, 10, 2), (4.6570, 5.6e-18, 2.5, 10, 3),
    (0.7600, 1.2e-17, 2.0, 16, 1), (1.7210, 7.0e-18, 2.5, 16, 2), (2.5690, 3.0e-18, 2.5, 16, 3),
]


def axes_symmetric(n, edge):
    """grid_mod.f90:533-548: real(i-1)/real(n-1) then * Rn, in FP32."""
    i = np.arange(n, dtype=np.float32)
    return (i / f32(n - 1)) * f32(edge)


def axes_full(n, edge):
    """grid_mod.f90:582-595: 2.*real(i-1)/real(n-1) - 1., then * Rn, in FP32."""
    i = np.arange(n, dtype=np.float32)
    return (f32(2.) * i / f32(n - 1) - f32(1.)) * f32(edge)


def active_map(x, y, z, R_out, R_in=0.0, dens=None):
    """setMotherGrid's active numbering (grid_mod.f90:1225-1294): loops
    i (x), j (y), k (z) with k fastest; radius = 1e10*sqrt(sum (a/1e10)^2)."""
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    s = f32(1.e10)
    r = s * np.sqrt((X / s) * (X / s) + (Y / s) * (Y / s) + (Z / s) * (Z / s))
    ok = ~((r < f32(R_in)) | ((r > f32(R_out)) & (f32(R_out) > 0)))
    if dens is not None:
        ok &= dens > 0
    flat = ok.reshape(-1)                      # C order = x outer, z inner
    act = np.zeros(flat.size, np.int32)
    act[flat] = np.arange(1, int(flat.sum()) + 1, dtype=np.int32)
    return np.asfortranarray(act.reshape(ok.shape)), int(flat.sum())


def locate_cell(a, v):
    """locate + midpoint rule of setStarPosition (grid_mod.f90:3590-3603), 1-based."""
    a = np.asarray(a, np.float32)
    v = f32(v)
    n = len(a)
    if v > a[-1]:
        i = n
    elif v < a[0]:
        i = 0
    else:
        d = a - v
        pos = np.nonzero(d > 0)[0]
        i = max(int(pos[np.argmin(d[pos])]) + 1 - 1, 1) if len(pos) else 1
    if i < n and v > (a[i - 1] + a[i]) / f32(2.):
        i += 1
    return max(i, 1)


def frequency_mesh(nbins, nu_min, nu_max, ion_edges):
    """initCartesianGrid's gas mesh (grid_mod.f90:213-266): the radio point,
    H series edges and ionisation thresholds (each +-0.0003 Ryd), then a log
    mesh from nuMin filling the remaining bins, sorted."""
    pts = []
    if nu_min < 1.489434e-6:
        pts.append(1.489434e-6)
    lo = nu_min
    for e in SERIES_EDGES + sorted(ion_edges):
        if e < nu_max:
            pts += [e, e - 0.0003, e + 0.0003]
            lo = min(lo, e - 0.0006)
    pts = [f32(p) for p in pts]
    n_log = nbins - len(pts)
    step = (np.log10(nu_max) - np.log10(lo)) / (n_log - 1)
    mesh = [lo * 10 ** (step * i) for i in range(n_log)]
    nu = np.sort(np.array(pts + mesh, np.float32))
    wid = np.empty_like(nu)
    wid[0] = nu[1] - nu[0]
    wid[1:-1] = (nu[2:] - nu[:-2]) / f32(2.)
    wid[-1] = nu[-1] - nu[-2]
    return nu, wid


def blackbody_cdf(nu, wid, T):
    """setProbDen (continuum_mod.f90:418-474) for a Planck spectrum."""
    x = HC_K_RYD * nu.astype(np.float64) / T
    with np.errstate(over="ignore"):
        bb = nu.astype(np.float64) ** 3 / np.expm1(np.minimum(x, 700.0))
    e = bb * wid
    cdf = np.cumsum(e) / e.sum()
    cdf = cdf.astype(np.float32)
    cdf[cdf >= cdf.max()] = 1.
    return cdf


def _sigma(nu, e):
    th, s0, p = e[0], e[1], e[2]
    out = np.where(nu >= th, s0 * (nu / th) ** (-p), 0.0)
    return out


def build_problem(n, edge, R_out, nH, Q_H, T_star, nbins, abundances, symmetric=True,
                  nu_min=1.e-5, nu_max=25., Te=1.e4, line_fraction=0.35, n_packets=1e8,
                  density=None, seed=0, nu=None):
    """Returns a dict with the grid (Fortran layouts) and tables of one
    synthetic problem.  abundances: {Z: X/H by number}; density: nH per
    lattice point [x, y, z] (default uniform nH); nu: a frequency mesh to use
    instead of the generated one (e.g. the one the host's atomic data are
    tabulated on)."""
    if symmetric:
        x = axes_symmetric(n, edge)
    else:
        x = axes_full(n, edge)
    y, z = x.copy(), x.copy()
    active, nc = active_map(x, y, z, R_out, dens=density)
    if nu is None:
        nu, wid = frequency_mesh(nbins, nu_min, nu_max, [e[0] for e in EDGES if e[3] in abundances])
    else:
        nu = np.asarray(nu, np.float32)
        nbins = nu.size
        wid = np.empty_like(nu)
        wid[0] = nu[1] - nu[0]
        wid[1:-1] = (nu[2:] - nu[:-2]) / f32(2.)
        wid[-1] = nu[-1] - nu[-2]
    cdf = blackbody_cdf(nu, wid, T_star)
    # source luminosity from Q(H): L = Q <h nu>_ion / (fraction of energy above 1 Ryd)
    nu64 = nu.astype(np.float64)
    xx = HC_K_RYD * nu64 / T_star
    with np.errstate(over="ignore"):
        bb = nu64 ** 3 / np.expm1(np.minimum(xx, 700.0)) * wid
    ion = nu64 >= 0.999434
    phot = (bb / nu64)[ion].sum()
    L_erg = Q_H * 2.1799e-11 * bb.sum() / phot               # erg/s
    L36 = L_erg / 1.e36
    # a model launched with no packets still has a finite packet energy
    deltaE = f32(L36 / max(n_packets, 1))

    # cell-centre radii and densities
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    r = np.sqrt(X.astype(np.float64) ** 2 + Y ** 2 + Z ** 2)
    cells = np.nonzero(active.reshape(-1, order="F") > 0)[0]
    ids = active.reshape(-1, order="F")[cells]
    rc = np.zeros(nc + 1)
    rc[ids] = r.reshape(-1, order="F")[cells]
    nden = np.zeros(nc + 1)
    nden[ids] = (density.reshape(-1, order="F")[cells] if density is not None else nH)

    # on-the-spot H ionisation along the radius: Q(r) drops by recombinations
    alphaB = 2.6e-13
    rr = np.linspace(0, max(rc.max(), 1.0) * 1.05, 4000)
    dr = rr[1] - rr[0]
    Qr = np.empty_like(rr)
    xi = np.empty_like(rr)
    q = Q_H
    n_mean = float(np.mean(nden[1:])) if nc > 0 else nH
    sig = 3.0e-18
    for i, rv in enumerate(rr):
        Qr[i] = q
        gam = q * sig / (4 * np.pi * max(rv, dr) ** 2)
        # x_ion^2 n alpha = (1-x_ion) gamma
        a_ = n_mean * alphaB
        disc = gam * gam + 4 * a_ * gam
        xion = (-gam + np.sqrt(disc)) / (2 * a_) if a_ > 0 else 1.0
        xion = min(max(xion, 1e-6), 1 - 1e-6)
        xi[i] = xion
        q = max(q - 4 * np.pi * rv * rv * dr * (n_mean * xion) ** 2 * alphaB, 0.0)
    xion_c = np.interp(rc, rr, xi)
    xH0 = np.clip(1 - xion_c, 1e-6, 1.0)
    xH0[0] = 0.0

    # tables, a block of cells at a time (elementwise in the cell index, so
    # the blocking does not change any value)
    opacity = np.zeros((nc + 1, nbins), np.float32, order="F")
    rec = np.zeros((nc + 1, nbins), np.float32, order="F")
    kT = Te / HC_K_RYD
    edges = [(e, _sigma(nu64, e)) for e in EDGES if e[3] in abundances]
    shapes = [nu64 ** 3 * sg * np.exp(-np.clip((nu64 - e[0]) / kT, 0, 700)) * (nu64 >= e[0])
              for e, sg in edges]
    ff_e = np.exp(-np.clip(nu64 / kT, 0, 700))
    ff_a = (1 - np.exp(-nu64 / kT)) * (1.0 / (nu64 * 3.28984e15) ** 3)
    def block(c0):
        sl = slice(c0, min(nc + 1, c0 + 16384))
        nd, x0 = nden[sl], xH0[sl]
        opac = np.zeros((nd.size, nbins))
        emis = np.zeros((nd.size, nbins))
        for (e, sg), shape in zip(edges, shapes):
            ab = abundances[e[3]]
            stage = e[4]
            # neutral stage follows H0; the next stages share the ionised part
            frac_lo = x0 if stage == 1 else (1 - x0) * (0.6 if stage == 2 else 0.3)
            opac += (nd * ab * frac_lo)[:, None] * sg[None, :]
            # Milne recombination continuum of the stage above
            rec_den = nd * ab * (1 - x0) * (1.0 if stage == 1 else 0.5)
            emis += (rec_den * nd * (1 - x0))[:, None] * shape[None, :] * 1e3
        # free-free emission and absorption (hydrogenic, Gaunt factor 1)
        ne = nd * (1 - x0)
        emis += (ne * ne)[:, None] * ff_e[None, :] * 1e-17
        opac += (ne * ne * 3.7e8 / np.sqrt(Te))[:, None] * ff_a[None, :]
        w = emis * wid[None, :]
        tot = w.sum(axis=1, keepdims=True)
        tot[tot == 0] = 1.0
        r = (np.cumsum(w, axis=1) / tot).astype(np.float32)
        r[r > f32(0.999998)] = 1.
        opacity[sl] = opac.astype(np.float32)
        rec[sl] = r

    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(8) as ex:
        list(ex.map(block, range(1, nc + 1, 16384)))
    totalLines = np.full(nc + 1, line_fraction, np.float32)
    totalLines[0] = 0.
    geo = (x[-1] - x[-2]) / f32(2.)
    if symmetric:
        sp = (0.0, 0.0, 0.0)
    else:
        sp = (0.0, 0.0, 0.0)
    sidx = tuple(locate_cell(a, 0.0) for a in (x, y, z))
    return dict(
        xAxis=x, yAxis=y, zAxis=z, active=active, nCells=nc, nbins=nbins, nuArray=nu,
        widFlx=wid, opacity=opacity, recPDF=rec, totalLines=totalLines,
        specCDF=np.asfortranarray(cdf.reshape(nbins, 1)),
        R_out=f32(R_out), geoCorr=(geo, geo, geo), ionEdge1=f32(0.999434), deltaE=deltaE,
        starPos=(0.0, 0.0, 0.0), starIdx=sidx, symmetric=symmetric, L36=L36, Hden=nden)


FULL_GAS = {1: 1.0, 2: 0.1, 6: 10 ** -3.6, 7: 1e-4, 8: 10 ** -3.3, 10: 1e-4, 16: 1e-5}
H_HE = {1: 1.0, 2: 0.1}


def lognormal_density(n, seed, mean=300.0, sigma_ln=0.5, corr=8.0, rng=None):
    """Lognormal nH field on an n^3 lattice [x, y, z]: Gaussian-filtered white
    noise (correlation length `corr` cells, periodic), scaled to standard
    deviation sigma_ln in ln nH, exp(mu + sigma_ln g) with
    mu = ln(mean) - sigma_ln^2 / 2 (so that the field's mean is `mean`)."""
    from scipy.ndimage import gaussian_filter
    rng = np.random.default_rng(seed) if rng is None else rng
    g = gaussian_filter(rng.standard_normal((n, n, n)), corr, mode="wrap")
    g = (g - g.mean()) / g.std()
    mu = np.log(mean) - 0.5 * sigma_ln ** 2
    return np.exp(mu + sigma_ln * g)


def clumpy_density(n, seed, mean=300.0, sigma_ln=0.5, corr=8.0, n_clumps=200, contrast=10.0,
                   clump_width=2.0):
    """The lognormal field of lognormal_density times `n_clumps` seeded
    Gaussian clumps of peak contrast `contrast` and width `clump_width` cells."""
    rng = np.random.default_rng(seed)
    dens = lognormal_density(n, seed, mean, sigma_ln, corr, rng=rng)
    w = clump_width
    r = int(np.ceil(4 * w))
    off = np.arange(-r, r + 1)
    bump = np.exp(-(off[:, None, None] ** 2 + off[None, :, None] ** 2 + off[None, None, :] ** 2) / (2 * w * w))
    boost = np.ones_like(dens)
    for c in rng.integers(0, n, (n_clumps, 3)):
        sl = tuple(slice(max(0, ci - r), min(n, ci + r + 1)) for ci in c)
        bs = tuple(slice(s_.start - (ci - r), s_.stop - (ci - r)) for s_, ci in zip(sl, c))
        boost[sl] += (contrast - 1.0) * bump[bs]
    return dens * boost


def workload(name, n_packets=1e8, seed=0, nu=None):
    """Named workloads (edge length, density, source and mesh of the decks).
    nu: frequency mesh to use (default: generated)."""
    if name == "c2":
        return build_problem(45, 2.5e18, 2.5e18, 1.e3, 1.e49, 4.e4, 800, FULL_GAS,
                             n_packets=n_packets, nu=nu)
    if name == "c1":
        return build_problem(19, 3.e18, 3.e18, 100., 4.e48, 4.e4, 400, H_HE,
                             n_packets=n_packets, nu=nu)
    if name == "c2small":
        return build_problem(15, 2.5e18, 2.5e18, 1.e3, 1.e49, 4.e4, 200, FULL_GAS,
                             n_packets=n_packets, nu=nu)
    if name == "c4":
        # clumpy 100^3 full grid, H+He gas and the 40,000 K, Q(H) = 4e48 source
        return build_problem(100, 3.e18, 3.e18, 300., 4.e48, 4.e4, 400, H_HE, symmetric=False,
                             n_packets=n_packets, density=clumpy_density(100, seed), nu=nu)
    raise KeyError(name)


def grain_tables(nu, Hden, path=None):
    """Grain tables of astronomical silicate (10 sizes 0.005-0.25 um, a^-3.5)
    on the frequency mesh nu, for timing the grain steps on a workload: the
    absorption cross-sections, emission integrals and grain constants the
    reference dumps for its dust deck (tests/golden/balance_dust_pe.npz),
    the cross-sections interpolated in log nu; Ndust in the deck's
    proportion to nH.  Returns the keyword arguments of
    Transport.dust_setup (photoelectric on)."""
    import os
    if path is None:
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "tests", "golden",
                            "balance_dust_pe.npz")
    g = dict(np.load(path))
    nS, nSz, nT, nsm, nSp = (int(v) for v in g["dust_ints"][:5])
    nu0 = np.asarray(g["nuArray"], np.float64)
    ab0 = g["dust_abs"].reshape(nS, nSz, -1).astype(np.float64)
    lnu = np.log(np.asarray(nu, np.float64))
    absx = np.empty((nS, nSz, lnu.size), np.float32)
    for s in range(nS):
        for ai in range(nSz):
            absx[s, ai] = np.exp(np.interp(lnu, np.log(nu0), np.log(np.maximum(ab0[s, ai], 1e-40))))
    Hd = np.asarray(g["Hden"], np.float64)
    nd = np.asarray(g["dust_ndust"], np.float64)
    ratio = float(np.median(nd[1:] / Hd[1:]))
    Ndust = (np.asarray(Hden, np.float64) * ratio).astype(np.float32)
    Ndust[0] = 0.0
    return dict(absXsec=absx, grainWeight=g["dust_weight"], grainAbun=g["dust_abun"],
                emIntegral=g["dust_emint"].reshape(nT, nSz, nSp).transpose(2, 1, 0), nTemps=nT,
                nSpeciesMax=nsm, photoelectric=True, grainRadius=g["dust_radius"], grainVn=g["dust_vn"],
                MsurfAtom=g["dust_msurf"], aWeight=g["dust_aweight"], Ndust=Ndust)
