"""Pins the oracle's spectral, BSDF and environment-light restatement.

Closed forms, published constants and independent numerical integrals, each
chosen so a plausible slip (a transposed table axis, a dropped Jacobian, a
missing 1/eta^2, a wrong normalisation) fails one of them.
"""
import math

import numpy as np
import pytest
from scipy import stats

import vision_amd as va
from oracle import oracle
from vision_amd import scenegen

pytestmark = pytest.mark.filterwarnings("ignore::RuntimeWarning")

RAW_CIE = None


def raw_cie():
    global RAW_CIE
    if RAW_CIE is None:
        RAW_CIE = oracle.data("cie_xyz_d65.f32", 4 * 471).reshape(4, 471).astype(np.float64)
    return RAW_CIE


XYZ2RGB = np.array([[3.240479, -1.537150, -0.498535], [-0.969256, 1.875991, 0.041556],
                    [0.055648, -0.204043, 1.057311]])


# --------------------------------------------------------------- CIE and wavelengths
def test_cie_y_integral_published_value():
    # CIE 1931 Y integrated over 360-830 nm is 106.856895 (pbrt's CIE_Y_integral);
    # the reference takes the 5 nm trapezoid (spd.cpp:9-17)
    spd, yi = oracle.cie_tables()
    assert abs(yi - 106.856895) < 0.05
    raw = raw_cie()
    assert spd[0] == np.float32(raw[0, 0]) and spd[94 + 20] == np.float32(raw[1, 100])
    assert spd[3 * 94 + 93] == np.float32(raw[3, 465])


def test_spd_eval_interpolates_its_samples():
    tab = np.arange(94, dtype=np.float32) ** 2
    interval = 471.0 / 94
    lam = (360 + interval * np.arange(0, 93)).astype(np.float32)
    got = oracle.spd_eval(tab, interval, lam)
    assert np.allclose(got, tab[:93], rtol=2e-5, atol=2e-3)
    mid = oracle.spd_eval(tab, interval, np.float32(360 + 10.5 * interval))
    assert abs(mid[0] - 0.5 * (tab[10] + tab[11])) < 1e-2
    # clamped outside 360-830
    assert oracle.spd_eval(tab, interval, np.float32(200.0))[0] == tab[0]


def test_wavelength_pdf_normalised_and_matches_sampler():
    lam = np.linspace(360, 830, 200001)
    pdf = oracle.wavelength_pdf(lam.astype(np.float32)).astype(np.float64)
    total = np.trapezoid(pdf, lam)
    assert abs(total - 1.0) < 2e-3
    rng = np.random.default_rng(7)
    u = rng.random(400_000, dtype=np.float32)
    lam_s, pdf_s = oracle.wavelengths(u)
    assert lam_s.min() > 359.5 and lam_s.max() < 830.5
    # the pdf returned with each sample is the pdf at that wavelength
    assert np.allclose(pdf_s, oracle.wavelength_pdf(lam_s))
    edges = np.linspace(360, 830, 48)
    counts, _ = np.histogram(lam_s, edges)
    fine = np.linspace(360, 830, 47 * 200 + 1)
    pf = oracle.wavelength_pdf(fine.astype(np.float32)).astype(np.float64)
    # the sampler covers [lambda(0), lambda(1)), slightly inside 360-830
    lo = oracle.wavelengths(np.array([0.0], np.float32))[0][0]
    hi = oracle.wavelengths(np.array([1.0 - 2 ** -24], np.float32))[0][0]
    pf[(fine < lo) | (fine > hi)] = 0
    cdf = np.concatenate([[0], np.cumsum(0.5 * (pf[1:] + pf[:-1]) * np.diff(fine))])
    exp = np.diff(cdf[::200]) / cdf[-1] * len(u)
    chi2 = np.sum((counts - exp) ** 2 / exp)
    assert stats.chi2.sf(chi2, len(exp) - 1) > 1e-3


def test_sigmoid_closed_form():
    x = np.array([-3.0, -0.5, 0.0, 0.25, 4.0], np.float32)
    coef = np.stack([np.zeros_like(x), np.zeros_like(x), x, np.ones_like(x)], 1)
    got = oracle.sigmoid(coef, np.array([500.0], np.float32))[:, 0]
    want = 0.5 + 0.5 * x / np.sqrt(1 + x.astype(np.float64) ** 2)
    assert np.allclose(got, want, atol=1e-6)
    inf = oracle.sigmoid(np.array([[0, 0, np.inf, 1], [0, 0, -np.inf, 1]], np.float32), np.array([600.0], np.float32))
    assert inf[0, 0] == 1.0 and inf[1, 0] == 0.0
    # the polynomial is c0 l^2 + c1 l + c2
    c = np.array([[1e-5, -0.01, 2.0, 1.0]], np.float32)
    lam = np.float64(550.0)
    xx = 1e-5 * lam * lam - 0.01 * lam + 2.0
    assert abs(oracle.sigmoid(c, np.array([550.0], np.float32))[0, 0] - (0.5 + 0.5 * xx / math.sqrt(1 + xx * xx))) < 1e-5


# --------------------------------------------------------------- RGB -> spectrum
def _reflectance_to_rgb(spectra_1nm):
    raw = raw_cie()
    xyz = spectra_1nm @ (raw[:3] * raw[3]).T / np.sum(raw[1] * raw[3])
    return xyz @ XYZ2RGB.T


def test_rgb_to_spectrum_round_trip():
    # the sigmoid spectrum of decode_albedo(rgb), lit by D65 and projected on
    # the CIE curves, reproduces rgb (the fit the coefficient table encodes)
    rng = np.random.default_rng(3)
    rgb = rng.uniform(0.05, 0.95, (300, 3)).astype(np.float32)
    rgb[:10] = rng.uniform(0.05, 0.95, (10, 1))  # greys take the closed-form branch
    coef = oracle.rgb_to_spectrum(rgb)
    lam = np.arange(360, 831, dtype=np.float32)
    spectra = oracle.sigmoid(coef, lam).astype(np.float64)
    back = _reflectance_to_rgb(spectra)
    err = np.abs(back - rgb)
    assert np.median(err) < 2e-3 and err.max() < 2e-2
    assert np.allclose(coef[:, 3], 0.212671 * rgb[:, 0] + 0.715160 * rgb[:, 1] + 0.072169 * rgb[:, 2], atol=1e-6)


def test_rgb_to_spectrum_unbound_scale():
    rgb = np.array([[2.0, 1.0, 0.5], [0.0, 0.0, 0.0], [4.0, 4.0, 4.0]], np.float32)
    c = oracle.rgb_to_spectrum(rgb, unbound=True)
    assert c[0, 3] == 4.0 and c[1, 3] == 0.0 and c[2, 3] == 8.0
    # a grey of 0.5 after scaling is the flat sigmoid 0.5: spectrum = scale / 2
    assert c[2, 0] == 0 and c[2, 1] == 0 and c[2, 2] == 0
    a = oracle.rgb_to_spectrum(rgb[:1] / 4.0)
    assert np.allclose(c[0, :3], a[0, :3])


def test_rgb_to_spectrum_product_matches_oracle_bitwise():
    rng = np.random.default_rng(11)
    rgb = np.concatenate([rng.uniform(-0.2, 1.3, (500, 3)), rng.uniform(0, 20, (100, 3)),
                          np.array([[0.3, 0.3, 0.3], [0, 0, 0], [1, 1, 1], [1, 0, 0], [0, 1e-6, 0]])]).astype(np.float32)
    for unbound in (False, True):
        want = oracle.rgb_to_spectrum(rgb, unbound)
        got = va.rgb_to_spectrum(rgb, unbound)
        assert np.array_equal(want.view(np.uint32), got.view(np.uint32)), unbound


# --------------------------------------------------------------- Fresnel and GGX
def test_fresnel_closed_forms():
    c = np.array([1.0, 0.0], np.float32)
    for eta in (1.33, 1.5, 2.4):
        f = oracle.fresnel(c, eta)
        assert abs(f[0] - ((eta - 1) / (eta + 1)) ** 2) < 1e-6 and abs(f[1] - 1.0) < 1e-6
    # total internal reflection from the dense side
    eta = 1 / 1.5
    crit = math.sqrt(1 - eta * eta)
    f = oracle.fresnel(np.array([crit * 0.9, crit * 1.1], np.float32), eta)
    assert f[0] == 1.0 and f[1] < 1.0
    # conductor at normal incidence and the k = 0 limit
    n, k = 0.2, 3.9
    fc = oracle.fresnel(np.array([1.0], np.float32), n, k)[0]
    assert abs(fc - ((n - 1) ** 2 + k * k) / ((n + 1) ** 2 + k * k)) < 1e-5
    cs = np.linspace(0.01, 1, 50).astype(np.float32)
    assert np.allclose(oracle.fresnel(cs, 1.5, 0.0), oracle.fresnel(cs, 1.5), atol=2e-6)


def _hemisphere_grid(nt, nphi, full=False):
    # midpoint rule in (cos theta, phi): equal solid-angle cells
    z = (np.arange(nt) + 0.5) / nt
    if full:
        z = z * 2 - 1
    phi = (np.arange(nphi) + 0.5) / nphi * 2 * np.pi
    Z, P = np.meshgrid(z, phi, indexing="ij")
    s = np.sqrt(1 - Z * Z)
    w = np.stack([s * np.cos(P), s * np.sin(P), Z], -1).reshape(-1, 3)
    dw = (2.0 if full else 1.0) / nt * 2 * np.pi / nphi
    return w.astype(np.float32), dw


@pytest.mark.parametrize("ax,ay", [(0.3, 0.3), (0.6, 0.6), (0.25, 0.7)])
def test_ggx_normalisation_and_visible_normals(ax, ay):
    wh, dw = _hemisphere_grid(1200, 256)
    wo = np.tile(np.array([[math.sin(0.6), 0.0, math.cos(0.6)]], np.float32), (len(wh), 1))
    D, G1, pdf = oracle.ggx(ax, ay, wo, wh)
    # int D(h) cos(theta_h) dh = 1
    assert abs(np.sum(D.astype(np.float64) * wh[:, 2]) * dw - 1) < 1e-2
    # the visible-normal density integrates to 1 over the normals facing wo
    # (weak white furnace); the sampler never produces back-facing normals
    front = (wh @ wo[0]) > 0
    assert abs(np.sum(pdf.astype(np.float64) * front) * dw - 1) < 1e-2
    # G1 is the Smith masking of GGX: 2 / (1 + sqrt(1 + alpha_o^2 tan^2))
    t2 = (math.tan(0.6)) ** 2
    a2 = ax * ax  # wo lies in the x-z plane
    assert abs(G1[0] - 2 / (1 + math.sqrt(1 + a2 * t2))) < 1e-5


def _chi2_sphere(samples, pdf_fn, n_z=24, n_phi=48, sub=6, n_total=None):
    """Chi-square of direction samples (on the full sphere) against pdf_fn;
    n_total counts draws including those rejected as invalid (pdf 0)."""
    z = np.clip(samples[:, 2], -1, 1)
    phi = np.mod(np.arctan2(samples[:, 1], samples[:, 0]), 2 * np.pi)
    iz = np.minimum(((z + 1) / 2 * n_z).astype(int), n_z - 1)
    ip = np.minimum((phi / (2 * np.pi) * n_phi).astype(int), n_phi - 1)
    counts = np.bincount(iz * n_phi + ip, minlength=n_z * n_phi).astype(np.float64)
    zz = -1 + (np.arange(n_z * sub) + 0.5) / (n_z * sub) * 2
    pp = (np.arange(n_phi * sub) + 0.5) / (n_phi * sub) * 2 * np.pi
    Z, P = np.meshgrid(zz, pp, indexing="ij")
    s = np.sqrt(1 - Z * Z)
    dirs = np.stack([s * np.cos(P), s * np.sin(P), Z], -1).reshape(-1, 3).astype(np.float32)
    dens = pdf_fn(dirs).astype(np.float64).reshape(n_z, sub, n_phi, sub).mean(axis=(1, 3)).ravel()
    cell = 2.0 / n_z * 2 * np.pi / n_phi
    expected = dens * cell * (len(samples) if n_total is None else n_total)
    keep = expected > 5
    chi2 = np.sum((counts[keep] - expected[keep]) ** 2 / expected[keep])
    # mass outside the kept cells must be small in both
    other = counts[~keep].sum(), expected[~keep].sum()
    return stats.chi2.sf(chi2, keep.sum() - 1), other


@pytest.mark.parametrize("ax,ay,theta", [(0.3, 0.3, 0.5), (0.5, 0.2, 1.2)])
def test_ggx_sampling_matches_visible_normal_pdf(ax, ay, theta):
    rng = np.random.default_rng(5)
    n = 300_000
    wo1 = np.array([math.sin(theta) * 0.8, math.sin(theta) * 0.6, math.cos(theta)], np.float32)
    wo = np.tile(wo1, (n, 1))
    wh = oracle.ggx_sample(ax, ay, wo, rng.random((n, 2), dtype=np.float32))
    # microfacet normals live in the upper hemisphere, facing wo
    p, other = _chi2_sphere(wh, lambda d: oracle.ggx(ax, ay, np.tile(wo1, (len(d), 1)), d)[2] *
                            ((d @ wo1) > 0) * (d[:, 2] > 0))
    assert p > 1e-3 and abs(other[0] - other[1]) < 0.01 * n


# --------------------------------------------------------------- BSDF records
def _rgb_spec():
    return oracle.Spectral(rgb=True)


def _mat_metal(rough, aniso=0.0, eta=(0.2, 0.9, 1.1), k=(3.9, 2.4, 2.2)):
    a = rough * rough
    if abs(aniso) <= 1e-4:
        ax = ay = a
    else:
        ax, ay = (a / (1 + aniso), a * (1 + aniso)) if aniso < 0 else (a * (1 - aniso), a / (1 - aniso))
    return oracle.material_record(3, [*eta, *k, ax, ay])


def _mat_glass(rough, ior=1.5, color=(1, 1, 1)):
    a = max(rough, 0.01) ** 2
    return oracle.material_record(4, [*color, 0.0, ior, a, a])


def _mat_principled(**kw):
    p = np.zeros(26, np.float32)
    p[0:3] = kw.get("color", (0.8, 0.5, 0.3))
    p[4:7] = kw.get("spec_tint", (1, 1, 1))
    p[8:11] = kw.get("sheen_tint", (1, 1, 1))
    p[12:15] = kw.get("coat_tint", (1, 1, 1))
    p[16] = kw.get("metallic", 0.3)
    p[17] = kw.get("ior", 1.5)
    p[18] = kw.get("roughness", 0.5)
    p[19] = kw.get("anisotropic", 0.0)
    p[20] = kw.get("sheen_weight", 0.3)
    p[21] = kw.get("sheen_roughness", 0.5)
    p[22] = kw.get("coat_weight", 0.3)
    p[23] = kw.get("coat_roughness", 0.2)
    p[24] = kw.get("coat_ior", 1.5)
    p[25] = kw.get("transmission_weight", 0.2)
    return oracle.material_record(5, p)


# eta of the transmitting (dielectric) lobe for the front side, else None
TRANSMIT_ETA = {"glass": 1.5, "principled": 1.5}


def _true_density_weight(name, wo1, d):
    """Weight turning the reported pdf (or f) into the sampler's density.

    The reference's PDF_wi_transmission (microfacet.h:150-156) leaves the
    eta^2 of dwh/dwi out and does not check that a direction is reachable by
    refraction through a front-facing microfacet; its sampler only reaches
    directions whose generalised half vector h (face-forwarded to wo) lies in
    the upper hemisphere with wi behind it.  Transmitted directions therefore
    carry eta^2 * [reachable] relative to the reported pdf."""
    eta = TRANSMIT_ETA.get(name)
    w = np.ones(len(d))
    if eta is None:
        return w
    h = wo1[None, :] + eta * d
    h /= np.linalg.norm(h, axis=1, keepdims=True)
    h = np.where((h @ wo1)[:, None] < 0, -h, h)
    reach = (h[:, 2] > 0) & (np.sum(d * h, 1) < 0)
    return np.where(d[:, 2] < 0, eta * eta * reach, 1.0)


MATERIALS = {
    "lambert": oracle.material_record(0, [0.7, 0.5, 0.3, 0.0]),
    "oren_nayar": oracle.material_record(1, [0.7, 0.5, 0.3, 0.0, 0.6]),
    "metal": _mat_metal(0.6),
    "metal_aniso": _mat_metal(0.7, aniso=0.5),
    "glass": _mat_glass(0.55),
    "principled": _mat_principled(),
    "principled_rough_metal": _mat_principled(metallic=0.9, roughness=0.8, anisotropic=0.4, transmission_weight=0.0),
}


@pytest.mark.parametrize("name", list(MATERIALS))
@pytest.mark.parametrize("theta", [0.4, 1.1])
def test_bsdf_sampling_matches_its_pdf(name, theta):
    # directions drawn by the material's sampler follow the pdf its
    # evaluator reports (the pair MIS relies on)
    spec = _rgb_spec()
    rec = MATERIALS[name]
    n = 300_000
    wo1 = np.array([math.sin(theta), 0.0, math.cos(theta)], np.float32)
    wo = np.tile(wo1, (n, 1))
    wi, f, pdf = oracle.bsdf_sample(spec, rec, wo, frame=3)
    valid = pdf > 0
    p, other = _chi2_sphere(wi[valid], lambda d: oracle.bsdf_eval(spec, rec, np.tile(wo1, (len(d), 1)), d)[1] *
                            _true_density_weight(name, wo1, d), n_total=n)
    assert p > 1e-3, (name, p)
    # the pdf reported with a valid sample equals the evaluator's at that direction
    k = np.flatnonzero(valid)[:2000]
    f2, pdf2 = oracle.bsdf_eval(spec, rec, wo[k], wi[k])
    assert np.allclose(pdf2, pdf[k], rtol=1e-5, atol=1e-7)
    assert np.allclose(f2, f[k], rtol=1e-5, atol=1e-7)


@pytest.mark.parametrize("name", list(MATERIALS))
def test_bsdf_estimators_agree(name):
    # E[f/pdf] over the sampler equals the integral of f|cos| by quadrature
    spec = _rgb_spec()
    rec = MATERIALS[name]
    wo1 = np.array([math.sin(0.7), 0.0, math.cos(0.7)], np.float32)
    n = 400_000
    wi, f, pdf = oracle.bsdf_sample(spec, rec, np.tile(wo1, (n, 1)), frame=9)
    est = np.where(pdf[:, None] > 0, f / np.maximum(pdf, 1e-30)[:, None], 0).astype(np.float64).mean(0)[:3]
    d, dw = _hemisphere_grid(900, 360, full=True)
    fq, _ = oracle.bsdf_eval(spec, rec, np.tile(wo1, (len(d), 1)), d)
    quad = (fq[:, :3].astype(np.float64) * _true_density_weight(name, wo1, d)[:, None]).sum(0) * dw
    assert np.allclose(est, quad, rtol=0.02, atol=2e-3), (est, quad)
    if name not in TRANSMIT_ETA:
        assert np.all(quad < 1.0 + 1e-3)


def _ggx_single_scatter_albedo(alpha, mu, nt=1500, nphi=512):
    """E_ss(alpha, mu) of an isotropic GGX mirror (F = 1) by quadrature of
    D G2 / (4 mu_o) over wi: textbook GGX, Smith height-correlated G2."""
    wi, dw = _hemisphere_grid(nt, nphi)
    wi = wi.astype(np.float64)
    wo = np.array([math.sqrt(1 - mu * mu), 0, mu])
    h = wi + wo
    h /= np.linalg.norm(h, axis=1, keepdims=True)
    c2 = h[:, 2] ** 2
    a2 = alpha * alpha
    D = a2 / (np.pi * ((a2 - 1) * c2 + 1) ** 2)

    def lam(w):
        t2 = (1 - w[..., 2] ** 2) / (w[..., 2] ** 2)
        return 0.5 * (np.sqrt(1 + a2 * t2) - 1)

    G2 = 1 / (1 + lam(wi) + lam(wo))
    return np.sum(D * G2 / (4 * mu)) * dw


@pytest.mark.parametrize("rough", [0.55, 0.8, 1.0])
@pytest.mark.parametrize("mu", [0.3, 0.8])
def test_conductor_energy_compensation_table(rough, mu):
    # ConductorLobe: f = f_ss / E(x = alpha, y = cos), the table holding the
    # single-scatter albedo of alpha' = x^2 (MicrofacetLobe::from_ratio_x);
    # so a white conductor reflects E_ss(alpha) / E_ss(alpha^2)
    alpha = rough * rough
    spec = _rgb_spec()
    rec = _mat_metal(rough, eta=(1.0, 1.0, 1.0), k=(1e4, 1e4, 1e4))
    wo1 = np.array([math.sqrt(1 - mu * mu), 0.0, mu], np.float32)
    d, dw = _hemisphere_grid(1500, 512)
    f, _ = oracle.bsdf_eval(spec, rec, np.tile(wo1, (len(d), 1)), d)
    E = f[:, 0].astype(np.float64).sum() * dw
    want = _ggx_single_scatter_albedo(alpha, mu) / _ggx_single_scatter_albedo(alpha * alpha, mu)
    assert abs(E - want) < 0.03 * want, (E, want)


@pytest.mark.parametrize("rough", [0.3, 0.6, 0.9])
@pytest.mark.parametrize("ior,mu", [(1.5, 0.9), (1.5, 0.4), (1 / 1.5, 0.9)])
def test_dielectric_white_furnace_with_compensation(rough, ior, mu):
    # DielectricLobe with kt = 1: reflection + transmission (in importance
    # transport, i.e. without the radiance 1/eta^2) divided by the table's
    # albedo is 1; the table is indexed by sqrt(alpha) like its precompute
    spec = _rgb_spec()
    rec = _mat_glass(rough, ior=ior)
    wo1 = np.array([math.sqrt(1 - mu * mu), 0.0, mu], np.float32)
    n = 400_000
    wi, f, pdf = oracle.bsdf_sample(spec, rec, np.tile(wo1, (n, 1)), frame=1)
    w = np.where(pdf > 0, f[:, 0] / np.maximum(pdf, 1e-30), 0).astype(np.float64)
    trans = wi[:, 2] < 0
    w[trans] *= ior * ior  # undo the radiance-transport 1/eta^2
    assert abs(w.mean() - 1.0) < 0.03, w.mean()


def test_lambert_and_oren_nayar_albedo():
    spec = _rgb_spec()
    wo1 = np.array([math.sin(0.5), 0, math.cos(0.5)], np.float32)
    d, dw = _hemisphere_grid(400, 200)
    f, pdf = oracle.bsdf_eval(spec, MATERIALS["lambert"], np.tile(wo1, (len(d), 1)), d)
    assert np.allclose(f[:, :3].astype(np.float64).sum(0) * dw, [0.7, 0.5, 0.3], rtol=1e-3)
    assert abs(pdf.astype(np.float64).sum() * dw - 1) < 1e-3
    f2, _ = oracle.bsdf_eval(spec, MATERIALS["oren_nayar"], np.tile(wo1, (len(d), 1)), d)
    alb = f2[:, :3].astype(np.float64).sum(0) * dw
    assert np.all(alb < [0.7, 0.5, 0.3])
    # seen from the normal, tan(beta) = 0 and Oren-Nayar is A R / pi with
    # A = 1 - s / (2 (s + 0.33)), s = (sigma pi/2)^4 as OrenNayar squares sigma^2
    # (bxdf.cpp:86-92)
    top = np.array([0, 0, 1], np.float32)
    f3, _ = oracle.bsdf_eval(spec, MATERIALS["oren_nayar"], np.tile(top, (len(d), 1)), d)
    sg = (0.6 * math.pi / 2) ** 4
    A = 1 - sg / (2 * (sg + 0.33))
    assert np.allclose(f3[:, :3].astype(np.float64).sum(0) * dw, A * np.array([0.7, 0.5, 0.3]), rtol=2e-3)


@pytest.mark.parametrize("kw", [dict(), dict(metallic=1.0, roughness=0.3), dict(sheen_weight=1.0, coat_weight=1.0),
                                dict(roughness=0.05, coat_weight=1.0, coat_roughness=0.05)])
def test_principled_white_energy_bound(kw):
    # a white principled stack without transmission reflects at most what
    # arrives (layering_weight keeps lower lobes under the upper ones'
    # albedo) and, with only opaque lobes, most of it
    spec = _rgb_spec()
    args = dict(color=(1, 1, 1), transmission_weight=0.0)
    args.update(kw)
    rec = _mat_principled(**args)
    for theta in (0.2, 0.9, 1.3):
        wo1 = np.array([math.sin(theta), 0, math.cos(theta)], np.float32)
        n = 200_000
        wi, f, pdf = oracle.bsdf_sample(spec, rec, np.tile(wo1, (n, 1)), frame=4)
        alb = np.where(pdf[:, None] > 0, f / np.maximum(pdf, 1e-30)[:, None], 0)[:, :3].astype(np.float64).mean(0)
        assert np.all(alb < 1.03) and np.all(alb > 0.6), (kw, theta, alb)


def test_glass_refraction_obeys_snell():
    spec = _rgb_spec()
    rec = _mat_glass(0.01, ior=1.5)  # alpha 1e-4: near-specular
    theta = 0.6
    wo1 = np.array([math.sin(theta), 0, math.cos(theta)], np.float32)
    wi, f, pdf = oracle.bsdf_sample(spec, rec, np.tile(wo1, (2000, 1)), frame=0)
    t = wi[:, 2] < 0
    sin_t = np.hypot(wi[t, 0], wi[t, 1])
    assert t.sum() > 1000
    assert np.allclose(sin_t, math.sin(theta) / 1.5, atol=2e-3)
    r = ~t & (pdf > 0)
    assert np.allclose(wi[r, 2], math.cos(theta), atol=2e-3)


# --------------------------------------------------------------- scene-level closed forms
def _hero(scene, dim=4):
    scene["spectrum"] = {"type": "hero", "param": {"dimension": dim}}
    return scene


def _mean_rgb(scene_dict, spp, base_dir="."):
    with va.Context(-1) as ctx:
        s = ctx.load_text(scene_dict, base_dir).build_accel()
        exp = s.export()
        film = oracle.render(exp, 0, spp)
        s.close()
    return film, exp


def _expected_white_xyz():
    """E[XYZ] of a D65 white over the hero sampler's wavelength range, from
    the interpolated CIE / D65 tables (SPD::eval) and the CIE Y integral."""
    spd, yi = oracle.cie_tables()
    l0 = oracle.wavelengths(np.array([0.0], np.float32))[0][0]
    l1 = oracle.wavelengths(np.array([1.0 - 2 ** -24], np.float32))[0][0]
    lam = np.linspace(l0, l1, 40001).astype(np.float32)
    interval = 471.0 / 94
    d65 = oracle.spd_eval(spd[282:376], interval, lam).astype(np.float64)
    xyz = [np.trapezoid(oracle.spd_eval(spd[94 * c:94 * c + 94], interval, lam) * d65, lam) for c in range(3)]
    return np.array(xyz) / yi


def test_hero_white_point():
    # a D65 white converts to XYZ with Y close to 1 and to linear sRGB near (1, 1, 1)
    xyz = _expected_white_xyz()
    rgb = XYZ2RGB @ xyz
    assert abs(xyz[1] - 1) < 5e-3 and np.allclose(rgb, 1, atol=0.02)


def test_hero_furnace_closed_form():
    # grey albedo 0.5 (sigmoid coefficients 0 -> reflectance 0.5 at every
    # wavelength) inside a white D65 emitter: L = XYZ_white / (1 - 0.5)
    sc = _hero(scenegen.furnace_box(albedo=(0.5, 0.5, 0.5), radiance=(1, 1, 1), width=16, height=16))
    film, exp = _mean_rgb(sc, 32)
    assert exp["spectrum_rgb"] == 0 and exp["spectrum_dim"] == 4
    got = film[..., :3].reshape(-1, 3).astype(np.float64).mean(0)
    want = XYZ2RGB @ (_expected_white_xyz() / (1 - 0.5))
    assert np.allclose(got, want, rtol=0.015), (got, want)


def _env_scene(tmp_path, img, albedo=0.5, hero=False, scale=1.0):
    scenegen.write_pfm(tmp_path / "env.pfm", img)
    sc = scenegen.lit_plane(scenegen.env_light("env.pfm", scale=scale), albedo=albedo, width=8, height=8)
    if hero:
        _hero(sc)
    return sc


@pytest.mark.parametrize("hero", [False, True])
def test_env_constant_radiance_closed_form(tmp_path, hero):
    # a plane under a constant environment L reflects albedo * L (the
    # environment's sampling pdf and the MIS weights sum to one)
    img = np.full((32, 64, 3), 0.8, np.float32)
    film, _ = _mean_rgb(_env_scene(tmp_path, img, hero=hero, scale=1.5), 64, str(tmp_path))
    got = film[..., :3].reshape(-1, 3).astype(np.float64).mean(0)
    want = 0.5 * 0.8 * 1.5
    if hero:
        want = want * (XYZ2RGB @ _expected_white_xyz())
    assert np.allclose(got, want, rtol=0.02), got


def test_env_map_orientation_and_importance(tmp_path):
    # a bright band in the upper rows of the image, dark elsewhere: the lit
    # plane (normal +y) integrates L cos / pi over the upper hemisphere of
    # the map as the importance sampler sees it; computed here by quadrature
    # of the texel lookup (nearest texel, v = theta / pi, theta from +y)
    h, w = 64, 128
    img = np.zeros((h, w, 3), np.float32)
    img[4:12, :, :] = 5.0
    img[12:28, 10:40, 1] = 2.0
    film, exp = _mean_rgb(_env_scene(tmp_path, img), 256, str(tmp_path))
    o2w = exp["env_o2w"]
    # closed-form mapping of the default orientation: map "up" (theta = 0) is
    # world +y, phi = 90 degrees is world +z
    assert np.allclose(o2w @ [0, 0, 1], [0, 1, 0], atol=1e-6)
    assert np.allclose(o2w @ [0, 1, 0], [0, 0, 1], atol=1e-6)
    assert np.allclose(exp["env_w2o"] @ o2w, np.eye(3), atol=1e-6)
    d, dw = _hemisphere_grid(2000, 512)
    world = d[:, [0, 2, 1]] * [1, 1, 1]  # hemisphere around +y
    world = np.stack([d[:, 0], d[:, 2], d[:, 1]], 1)
    local = world @ exp["env_w2o"].T
    theta = np.arccos(np.clip(local[:, 2], -1, 1))
    phi = np.mod(np.arctan2(local[:, 1], local[:, 0]), 2 * np.pi)
    x = np.minimum((phi / (2 * np.pi) * w).astype(int), w - 1)
    y = np.minimum((theta / np.pi * h).astype(int), h - 1)
    L = img[y, x].astype(np.float64)
    want = 0.5 / np.pi * np.sum(L * world[:, 1:2], 0) * dw
    got = film[..., :3].reshape(-1, 3).astype(np.float64).mean(0)
    assert np.allclose(got, want, rtol=0.03, atol=2e-3), (got, want)


def _mat_mirror(rough, color=(1, 1, 1)):
    a = rough * rough
    return oracle.material_record(6, [*color, 0.0, 0.0, 0.0, a, a])


@pytest.mark.parametrize("rough", [0.55, 0.8, 1.0])
@pytest.mark.parametrize("mu", [0.3, 0.8])
def test_mirror_reflects_its_tint_with_compensation(rough, mu):
    # MirrorLobe: constant Fresnel, f = kr f_ss / E(alpha, cos); the
    # reflectance is kr E_ss(alpha) / E_ss(alpha^2), linear in the tint
    alpha = rough * rough
    spec = _rgb_spec()
    tint = np.array([0.9, 0.5, 0.2])
    wo1 = np.array([math.sqrt(1 - mu * mu), 0.0, mu], np.float32)
    d, dw = _hemisphere_grid(1500, 512)
    f, pdf = oracle.bsdf_eval(spec, _mat_mirror(rough, tint), np.tile(wo1, (len(d), 1)), d)
    fw, pdfw = oracle.bsdf_eval(spec, _mat_mirror(rough), np.tile(wo1, (len(d), 1)), d)
    np.testing.assert_allclose(f[:, :3], fw[:, :3] * tint.astype(np.float32), rtol=1e-6, atol=1e-12)
    np.testing.assert_array_equal(pdf, pdfw)
    E = f[:, :3].astype(np.float64).sum(axis=0) * dw
    want = tint * _ggx_single_scatter_albedo(alpha, mu) / _ggx_single_scatter_albedo(alpha * alpha, mu)
    np.testing.assert_allclose(E, want, rtol=0.03)
    # the white mirror equals a conductor whose Fresnel is 1 to within its
    # residual (eta 1, k 1e4: F = 1 - 4e-8)
    fc, _ = oracle.bsdf_eval(spec, _mat_metal(rough, eta=(1.0, 1.0, 1.0), k=(1e4, 1e4, 1e4)),
                             np.tile(wo1, (len(d), 1)), d)
    np.testing.assert_allclose(fw[:, :3], fc[:, :3], rtol=1e-5, atol=1e-12)


# ------------------------------------------------------------ principled stack reductions
# PrincipledBSDF::create_lobe_set (principled_bsdf.cpp:352-461) stacks sheen,
# coat, metallic, transmission, specular and diffuse lobes with
# layering_weight between them.  Each test below switches the stack down to
# one lobe whose closed form is known independently, so a wrong weight, a
# transposed table axis or a wrong Fresnel term fails it.

def _f32sq(r):
    r = np.float32(r)
    return float(r * r)


def _mat_mirror_alpha(alpha, color=(1, 1, 1)):
    return oracle.material_record(6, [*color, 0.0, 0.0, 0.0, alpha, alpha])


def _upper_pairs(n, seed):
    rng = np.random.default_rng(seed)

    def dirs():
        z = rng.uniform(0.05, 1.0, n)
        phi = rng.uniform(0, 2 * np.pi, n)
        s = np.sqrt(1 - z * z)
        return np.stack([s * np.cos(phi), s * np.sin(phi), z], 1).astype(np.float32)

    return dirs(), dirs()


def _metal_only(color, tint, rough):
    return _mat_principled(color=color, spec_tint=tint, metallic=1.0, roughness=rough, sheen_weight=0.0,
                           coat_weight=0.0, transmission_weight=0.0)


@pytest.mark.parametrize("rough", [0.3, 0.7])
def test_principled_metallic_white_is_the_compensated_ggx_mirror(rough):
    # metallic = 1, everything else 0: one MetallicLobe remains; white colour
    # and tint give F82 = 1, so the stack equals the white mirror (F = 1,
    # same GGX, same pure-reflection compensation), pdf included
    spec = _rgb_spec()
    wo, wi = _upper_pairs(20000, 1)
    f, pdf = oracle.bsdf_eval(spec, _metal_only((1, 1, 1), (1, 1, 1), rough), wo, wi)
    fm, pdfm = oracle.bsdf_eval(spec, _mat_mirror_alpha(_f32sq(rough)), wo, wi)
    assert fm[:, :3].max() > 0
    np.testing.assert_allclose(f[:, :3], fm[:, :3], rtol=2e-6, atol=1e-12)
    np.testing.assert_allclose(pdf, pdfm, rtol=2e-6, atol=1e-12)


def _f82_ratio(color, tint, rough, wo, wi):
    spec = _rgb_spec()
    f, _ = oracle.bsdf_eval(spec, _metal_only(color, tint, rough), wo, wi)
    fm, _ = oracle.bsdf_eval(spec, _mat_mirror_alpha(_f32sq(rough)), wo, wi)
    keep = fm[:, 0] > 1e-6
    h = wo.astype(np.float64) + wi.astype(np.float64)
    h /= np.linalg.norm(h, axis=1, keepdims=True)
    cos_oh = np.abs(np.sum(wo * h, 1))
    return (f[keep, :3] / fm[keep, :3]).astype(np.float64), cos_oh[keep]


@pytest.mark.parametrize("color", [(0.9, 0.6, 0.2), (0.05, 0.3, 0.8)])
def test_principled_f82_at_unit_tint_is_schlick(color):
    # FresnelF82Tint (fresnel.h:98-130) with F82 = 1 has B = 0: Schlick's
    # F0 + (1 - F0)(1 - cos)^5 with F0 = the base colour
    wo, wi = _upper_pairs(20000, 2)
    ratio, c = _f82_ratio(color, (1, 1, 1), 0.5, wo, wi)
    c0 = np.array(color)
    want = c0[None, :] + (1 - c0[None, :]) * (1 - c[:, None]) ** 5
    np.testing.assert_allclose(ratio, want, rtol=3e-5)


@pytest.mark.parametrize("color,tint", [((0.9, 0.6, 0.2), (0.5, 0.7, 0.9)), ((0.2, 0.4, 0.95), (1.0, 0.3, 0.6))])
def test_principled_f82_tint_at_cos_one_seventh(color, tint):
    # by construction F82(cos = 1/7) = tint * Schlick(1/7); F82(1) = colour
    c0, t0 = np.array(color), np.array(tint)
    schlick = lambda c: c0 + (1 - c0) * (1 - c) ** 5
    for cos_target, want in ((1 / 7, t0 * schlick(1 / 7)), (0.999, None)):
        s = math.sqrt(1 - cos_target * cos_target)
        # wh = +z, wo and wi mirror images: |wo . wh| = cos_target
        wo = np.array([[s, 0.0, cos_target]], np.float32)
        wi = np.array([[-s, 0.0, cos_target]], np.float32)
        ratio, c = _f82_ratio(color, tint, 0.5, wo, wi)
        if want is None:
            # near normal incidence the tint term vanishes as (1 - cos)^6
            np.testing.assert_allclose(ratio[0], c0, rtol=1e-4)
        else:
            assert abs(c[0] - 1 / 7) < 1e-6
            np.testing.assert_allclose(ratio[0], want, rtol=2e-5)


def _sheen_table():
    return oracle.data("lut_sheen_approx.f32", 4096).reshape(32, 32, 4).astype(np.float64)


@pytest.mark.parametrize("i,j", [(6, 20), (20, 6), (27, 11), (11, 27)])
def test_principled_sheen_alone_reflects_the_ltc_albedo(i, j):
    # sheen only (colour 0 turns off diffuse and transmission; ior 1 gives
    # the specular lobe zero Fresnel): its LTC is normalised over the upper
    # hemisphere, so f|cos| integrates to tint * R with R the table's third
    # channel at (cos theta_o, sheen roughness); the table's fast axis is
    # cos theta (principled_bsdf.cpp:49-51, make_float2(cos_theta, alpha)).
    # At texel centres the bilinear lookup returns the entry itself.
    tab = _sheen_table()
    cos_o, rough = (i + 0.5) / 32, (j + 0.5) / 32
    R = tab[j, i, 2]
    assert abs(R - tab[i, j, 2]) > 1e-3 * R  # a transposed lookup would read another value
    tint = np.array([0.9, 0.6, 0.3])
    rec = _mat_principled(color=(0, 0, 0), metallic=0.0, ior=1.0, sheen_weight=1.0, sheen_roughness=rough,
                          sheen_tint=tuple(tint), coat_weight=0.0, transmission_weight=0.0)
    wo1 = np.array([math.sqrt(1 - cos_o * cos_o), 0.0, cos_o], np.float32)
    d, dw = _hemisphere_grid(1200, 600)
    f, _ = oracle.bsdf_eval(_rgb_spec(), rec, np.tile(wo1, (len(d), 1)), d)
    E = f[:, :3].astype(np.float64).sum(0) * dw
    np.testing.assert_allclose(E, tint * R, rtol=0.01)


def _coat_only(tint, rough, ior):
    return _mat_principled(color=(0, 0, 0), metallic=0.0, ior=1.0, sheen_weight=0.0, coat_weight=1.0,
                           coat_roughness=rough, coat_ior=ior, coat_tint=tint, transmission_weight=0.0)


@pytest.mark.parametrize("rough,ior", [(0.3, 1.5), (0.6, 2.2)])
def test_principled_coat_alone_is_a_dielectric_ggx_reflection(rough, ior):
    # coat only: MicrofacetReflection with the exact dielectric Fresnel of
    # the coat ior, GGX alpha = roughness^2, tinted, no compensation:
    # f |cos i| = tint D G2 F / (4 cos o), textbook forms in float64
    tint = np.array([0.8, 0.5, 0.9])
    wo, wi = _upper_pairs(20000, 3)
    f, _ = oracle.bsdf_eval(_rgb_spec(), _coat_only(tuple(tint), rough, ior), wo, wi)
    a = _f32sq(rough)
    a2 = a * a
    o, i = wo.astype(np.float64), wi.astype(np.float64)
    h = o + i
    h /= np.linalg.norm(h, axis=1, keepdims=True)
    D = a2 / (np.pi * ((a2 - 1) * h[:, 2] ** 2 + 1) ** 2)
    lam = lambda w: 0.5 * (np.sqrt(1 + a2 * (1 - w[:, 2] ** 2) / w[:, 2] ** 2) - 1)
    G2 = 1 / (1 + lam(o) + lam(i))
    c = np.sum(o * h, 1)
    g = np.sqrt(ior * ior - 1 + c * c)
    F = 0.5 * ((g - c) / (g + c)) ** 2 * (1 + ((c * (g + c) - 1) / (c * (g - c) + 1)) ** 2)
    want = tint[None, :] * (D * G2 * F / (4 * o[:, 2]))[:, None]
    np.testing.assert_allclose(f[:, :3], want, rtol=2e-4, atol=1e-7)


def test_principled_zero_coat_weight_leaves_the_stack_unchanged():
    # with coat_weight 0 the coat's roughness, ior and tint have no effect
    spec = _rgb_spec()
    wo, wi = _upper_pairs(5000, 4)
    a = _mat_principled(coat_weight=0.0, coat_roughness=0.2, coat_ior=1.5, coat_tint=(1, 1, 1))
    b = _mat_principled(coat_weight=0.0, coat_roughness=0.8, coat_ior=2.7, coat_tint=(0.1, 0.9, 0.3))
    fa, pa = oracle.bsdf_eval(spec, a, wo, wi)
    fb, pb = oracle.bsdf_eval(spec, b, wo, wi)
    np.testing.assert_array_equal(fa, fb)
    np.testing.assert_array_equal(pa, pb)
    # and a nonzero coat does change it
    fc, _ = oracle.bsdf_eval(spec, _mat_principled(coat_weight=0.5), wo, wi)
    assert np.abs(fc - fa).max() > 1e-3


@pytest.mark.parametrize("rough,ior", [(0.3, 1.5), (0.6, 1.33)])
def test_principled_full_transmission_is_the_rough_dielectric(rough, ior):
    # transmission 1 (metallic, sheen, coat 0): the weight below it is 0, so
    # only the DielectricLobe remains; with spec_tint 1 its Schlick Fresnel
    # F0 + t (1 - F0), t = (F - F0) / (1 - F0), is the exact dielectric F,
    # so the stack equals the glass material of the same colour, ior and
    # alpha, over reflection and transmission, entering from either side
    # (inside, the lobe's eta and Schlick F0 use 1 / ior)
    spec = _rgb_spec()
    color = (0.9, 0.7, 0.5)
    rec = _mat_principled(color=color, spec_tint=(1, 1, 1), metallic=0.0, sheen_weight=0.0, coat_weight=0.0,
                          transmission_weight=1.0, roughness=rough, ior=ior)
    a = _f32sq(rough)
    glass = oracle.material_record(4, [*color, 0.0, ior, a, a])
    rng = np.random.default_rng(5)
    n = 20000
    wo, _ = _upper_pairs(n, 6)
    wo[n // 2:, 2] *= -1
    z = rng.uniform(-1, 1, n)
    phi = rng.uniform(0, 2 * np.pi, n)
    s = np.sqrt(1 - z * z)
    wi = np.stack([s * np.cos(phi), s * np.sin(phi), z], 1).astype(np.float32)
    f, pdf = oracle.bsdf_eval(spec, rec, wo, wi)
    fg, pdfg = oracle.bsdf_eval(spec, glass, wo, wi)
    assert (fg[:, 0] > 0).sum() > n // 2 and (fg[wi[:, 2] * wo[:, 2] < 0, 0] > 0).any()
    assert (fg[n // 2:, 0] > 0).sum() > n // 8
    np.testing.assert_allclose(f[:, :3], fg[:, :3], rtol=2e-5, atol=1e-9)
    np.testing.assert_allclose(pdf, pdfg, rtol=2e-5, atol=1e-9)


# ------------------------------------------------------------ mix / add
# MixMaterial / AddMaterial (mix.cpp, add.cpp) through LobeSet::create_mix /
# create_add and flatten (lobe.cpp:495-562).  A mix is the (1 - frac, frac)
# blend of its two components, in f and in pdf; an add sums the
# evaluations of single-lobe components, halves those of lobe-set
# components (flatten scales their lobes by the normalised sample weight
# 1/2) and halves every pdf.

def _mat_mix(frac, a, b, add=False):
    rec = np.zeros(32, np.uint32)
    rec[0] = 8 if add else 7
    rec[2], rec[3] = 1, 2
    rec[4] = np.float32(frac).view(np.uint32)
    return np.concatenate([rec, np.asarray(a, np.uint32).reshape(32), np.asarray(b, np.uint32).reshape(32)])


def _sphere_pairs(n, seed):
    rng = np.random.default_rng(seed)
    wo, _ = _upper_pairs(n, seed)
    v = rng.normal(size=(n, 3))
    wi = (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)
    return wo, wi


@pytest.mark.parametrize("a,b", [("lambert", "metal"), ("principled", "oren_nayar"), ("glass", "principled"),
                                 ("metal_aniso", "glass")])
@pytest.mark.parametrize("frac", [0.0, 0.3, 1.0])
def test_mix_is_the_weighted_blend_of_its_components(a, b, frac):
    spec = _rgb_spec()
    wo, wi = _sphere_pairs(20000, 11)
    fa, pa = oracle.bsdf_eval(spec, MATERIALS[a], wo, wi)
    fb, pb = oracle.bsdf_eval(spec, MATERIALS[b], wo, wi)
    fm, pm = oracle.bsdf_eval(spec, _mat_mix(frac, MATERIALS[a], MATERIALS[b]), wo, wi)
    w0, w1 = np.float32(1 - np.float32(frac)), np.float32(frac)
    if frac == 0.0:
        np.testing.assert_allclose(fm, fa, rtol=1e-6, atol=1e-12)
    if frac == 1.0:
        np.testing.assert_allclose(fm, fb, rtol=1e-6, atol=1e-12)
    want_f = w0 * fa.astype(np.float64) + w1 * fb.astype(np.float64)
    want_p = w0 * pa.astype(np.float64) + w1 * pb.astype(np.float64)
    np.testing.assert_allclose(fm, want_f, rtol=2e-5, atol=1e-9)
    np.testing.assert_allclose(pm, want_p, rtol=2e-5, atol=1e-9)
    assert fm[:, :3].max() > 0


@pytest.mark.parametrize("a,b", [("lambert", "metal"), ("principled", "oren_nayar"), ("glass", "principled_rough_metal")])
def test_add_sums_single_lobes_and_halves_lobe_sets(a, b):
    spec = _rgb_spec()
    wo, wi = _sphere_pairs(20000, 12)
    fa, pa = oracle.bsdf_eval(spec, MATERIALS[a], wo, wi)
    fb, pb = oracle.bsdf_eval(spec, MATERIALS[b], wo, wi)
    fm, pm = oracle.bsdf_eval(spec, _mat_mix(0.0, MATERIALS[a], MATERIALS[b], add=True), wo, wi)
    ea = 0.5 if a.startswith("principled") else 1.0
    eb = 0.5 if b.startswith("principled") else 1.0
    np.testing.assert_allclose(fm, ea * fa.astype(np.float64) + eb * fb.astype(np.float64), rtol=2e-5, atol=1e-9)
    np.testing.assert_allclose(pm, 0.5 * pa.astype(np.float64) + 0.5 * pb.astype(np.float64), rtol=2e-5, atol=1e-9)


MIXES = {
    "mix_lambert_metal": _mat_mix(0.35, MATERIALS["lambert"], MATERIALS["metal"]),
    "mix_principled_oren": _mat_mix(0.6, MATERIALS["principled_rough_metal"], MATERIALS["oren_nayar"]),
    "add_metal_principled": _mat_mix(0.0, MATERIALS["metal_aniso"], MATERIALS["principled_rough_metal"], add=True),
}


@pytest.mark.parametrize("name", list(MIXES))
def test_mix_sampling_matches_its_pdf(name):
    # the flattened lobe set's sampler (one lobe picked by its sample
    # weight) follows the pdf its evaluator reports
    spec = _rgb_spec()
    rec = MIXES[name]
    n = 300_000
    theta = 0.8
    wo1 = np.array([math.sin(theta), 0.0, math.cos(theta)], np.float32)
    wo = np.tile(wo1, (n, 1))
    wi, f, pdf = oracle.bsdf_sample(spec, rec, wo, frame=5)
    valid = pdf > 0
    p, _ = _chi2_sphere(wi[valid], lambda d: oracle.bsdf_eval(spec, rec, np.tile(wo1, (len(d), 1)), d)[1], n_total=n)
    assert p > 1e-3, (name, p)
    k = np.flatnonzero(valid)[:2000]
    f2, pdf2 = oracle.bsdf_eval(spec, rec, wo[k], wi[k])
    np.testing.assert_allclose(pdf2, pdf[k], rtol=1e-5, atol=1e-7)
    np.testing.assert_allclose(f2, f[k], rtol=1e-5, atol=1e-7)
