#!/usr/bin/env python3
"""Write the synthetic spectral files and McICA data file used by the tests
and the benchmark into socrates_amd/data/.

Run once in the build container (it reads data blocks of the reference's
dev spectral files as *data*: band limits, solar spectrum, Rayleigh
coefficients, Planck polynomial, droplet and ice parametrisations). The gas
k-distributions are synthetic, drawn with a fixed seed:

  * 9 LW bands / 6 SW bands, one major gas per band with 1-8 k-terms
    (50 LW / 40 SW g-points), k log-uniform in 1e-6..1e2 m2/kg, weights
    normalised to 1 per band, power-law scaling (ip_scale_term,
    ip_scale_power_law) with p_ref=1e5 Pa, T_ref=250 K, exponents
    a ~ U(0.5,1), b ~ U(0,2);
  * 0-2 minor gases per band with 1-4 k-terms drawn the same way, so the
    equivalent-extinction path of solve_band_k_eqv_scl is exercised.

The files follow the text formats read by read_spectrum.F90 (block readers
1_0_0, 2_0_0, 3_0_0, 4_0_0, 5_0_1, 6_0_0, 10_0_2, 12_0_2) and by
read_mcica_data (def_mcica.F90:61-190). The McICA xcw table is synthetic:
quantiles of a gamma distribution with unit mean.

Usage: python3 tests/golden/make_spectra.py [--ref /root/reference]
"""
import argparse
import os

import numpy as np
from scipy import stats

REPO = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
OUT = os.path.join(REPO, "socrates_amd", "data")

GAS_NAMES = {1: "Water Vapour", 2: "Carbon Dioxide", 3: "Ozone"}

# (major gas, n_major_terms, [(minor gas, n_terms), ...]) per band
LW_BANDS = [
    (1, 6, []),
    (1, 7, [(2, 2)]),
    (2, 8, [(1, 3), (3, 1)]),
    (1, 5, [(3, 2)]),
    (3, 6, [(1, 2)]),
    (1, 4, [(2, 1)]),
    (1, 5, []),
    (1, 4, [(2, 2)]),
    (1, 5, []),
]
SW_BANDS = [
    (3, 4, [(1, 1)]),
    (1, 6, [(3, 2)]),
    (1, 8, []),
    (1, 8, [(2, 3)]),
    (1, 7, [(2, 2)]),
    (1, 7, []),
]


def read_block(path, btype):
    """Return the text lines (header .. *END inclusive) of block `btype`."""
    lines = open(path).read().splitlines()
    for i, ln in enumerate(lines):
        if ln.startswith("*BLOCK") and int(ln[15:19]) == btype:
            j = i
            while not lines[j].startswith("*END"):
                j += 1
            return lines[i:j + 1]
    raise KeyError(btype)


def e16(x):
    return "%16.9E" % x


def block0(n_band, gases):
    out = ["*BLOCK: TYPE =    0: SUBTYPE =    0: VERSION =    1",
           "Summary of spectral data.",
           "Number of spectral bands = %5d" % n_band,
           "Total number of gaseous absorbers = %5d" % len(gases),
           "Total number of aerosols = %5d" % 0,
           "List of indexing numbers and absorbers.",
           "Index       Absorber(identifier and name)"]
    for i, g in enumerate(gases):
        out.append("%5d       %5d       %s" % (i + 1, g, GAS_NAMES[g]))
    out += ["List of indexing numbers of aerosols.",
            "Index       Aerosol(type number and name)", "*END"]
    return out


def block4(bands, gas_index):
    out = ["*BLOCK: TYPE =    4: SUBTYPE =    0: VERSION =    0",
           "Gaseous absorbers in each interval",
           "(The number is the indexing number of the species as set out",
           " in the summary block 0.)",
           "A zero indicates that there in no gaseous absorption in the interval.",
           "Band        Number of active absorbers followed by indexing numbers"]
    for ib, (major, _, minors) in enumerate(bands):
        ids = [gas_index[major]] + [gas_index[g] for g, _ in minors]
        out.append("%5d       %5d" % (ib + 1, len(ids)))
        out.append("     " + "".join("  %3d" % i for i in ids))
    out.append("*END")
    return out


def k_terms(rng, n):
    k = 10.0 ** rng.uniform(-6.0, 2.0, n)
    k.sort()
    w = rng.uniform(0.05, 1.0, n)
    w /= w.sum()
    # renormalise so that the weights sum to 1 to rounding of the printed digits
    w = np.array([float("%16.9E" % x) for x in w])
    w[-1] = 1.0 - w[:-1].sum()
    a = rng.uniform(0.5, 1.0, n)
    b = rng.uniform(0.0, 2.0, n)
    return k, w, a, b


def block5(rng, bands, gas_index):
    out = ["*BLOCK: TYPE =    5: SUBTYPE =    0: VERSION =    1",
           "Exponential sum fiting coefficients: (exponents: m2/kg)",
           "Band        Gas, Number of k-terms, Scaling type and scaling function,",
           "             followed by reference pressure and temperature,",
           "                   k-terms, weights and scaling parameters."]
    for ib, (major, nmaj, minors) in enumerate(bands):
        for g, n in [(major, nmaj)] + list(minors):
            out.append("%5d       %5d       %5d       %5d       %5d"
                       % (ib + 1, gas_index[g], n, 2, 1))
            out.append("      %s      %s" % (e16(1.0e5), e16(250.0)))
            k, w, a, b = k_terms(rng, n)
            for i in range(n):
                out.append("   %s   %s%3d%s   %s" % (e16(k[i]), e16(w[i]), 0,
                                                     e16(a[i]), e16(b[i])))
    out.append("*END")
    return out


def block8(n_band):
    """No continuum absorption in any band (read_block_8_0_0)."""
    out = ["*BLOCK: TYPE =    8: SUBTYPE =    0: VERSION =    0",
           "Continuum absorbers in each interval",
           "(The number is the indexing number of each type as set out",
           " in the module rad_pcf.)",
           "A zero indicates that there is no continuum absorption in the interval.",
           "Band         Number of active continua followed by indexing numbers"]
    for b in range(n_band):
        out.append("%5d       %5d" % (b + 1, 0))
    out += ["Indexing numbers of gases for continua:",
            "     Index of water = %5d" % 1, "*END"]
    return out


def block6_from_dev(dev_lines):
    """Planck polynomial: re-emit the dev coefficients in the 6_0_0 layout."""
    deg, tref, coeffs = None, None, []
    for ln in dev_lines:
        if ln.startswith("Degree of polynomial"):
            deg = int(ln.split("=")[1].split(":")[0])
            tref = float(ln.split("=")[2])
    vals = []
    start = [i for i, ln in enumerate(dev_lines) if ln.startswith("Band")][0] + 1
    for ln in dev_lines[start:-1]:
        vals += [float(x) for x in ln.split()]
    nb = len(vals) // (deg + 2)
    for b in range(nb):
        row = vals[b * (deg + 2):(b + 1) * (deg + 2)]
        coeffs.append(row[1:])
    out = ["*BLOCK: TYPE =    6: SUBTYPE =    0: VERSION =    0",
           "Thermal source function.",
           "Degree of polynomial = %5d: Reference temperature = %s" % (deg, e16(tref)),
           "Coefficients in each band.",
           "Band          w/m2"]
    for b, c in enumerate(coeffs):
        first = "%5d       " % (b + 1)
        for j in range(0, len(c), 3):
            chunk = "".join("%s    " % e16(x) for x in c[j:j + 3])
            out.append((first if j == 0 else " " * 12) + chunk.rstrip())
    out.append("*END")
    return out


def write_spectrum(path, dev, bands, rng, solar):
    gases = sorted({g for b in bands for g in [b[0]] + [m for m, _ in b[2]]})
    gas_index = {g: i + 1 for i, g in enumerate(gases)}
    lines = block0(len(bands), gases)
    lines += read_block(dev, 1)
    if solar:
        lines += read_block(dev, 2)
        lines += read_block(dev, 3)
    lines += block4(bands, gas_index)
    lines += block5(rng, bands, gas_index)
    lines += block8(len(bands))
    if not solar:
        lines += block6_from_dev(read_block(dev, 6))
    lines += read_block(dev, 10)
    lines += read_block(dev, 12)
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def write_mcica(path, rng, lw_bands, sw_bands):
    n_gen = 64
    # two sampled sub-columns for a few g-points so that subcol_k > 1 is covered
    lw_k2 = {(3, 1), (6, 2)}
    sw_k2 = {(4, 2)}
    n_lw = sum(n for _, n, _ in lw_bands) + len(lw_k2)
    n_sw = sum(n for _, n, _ in sw_bands) + len(sw_k2)
    n_req = max(n_lw, n_sw)
    reorder = rng.permutation(np.arange(1, n_req + 1))
    n1, n2 = 100, 40
    p = (np.arange(n1) + 0.5) / n1
    xcw = np.zeros((n1, n2))
    for j in range(n2):
        rsd = (j + 1 + 3.0) / 40.0
        shape = 1.0 / rsd ** 2
        xcw[:, j] = stats.gamma.ppf(p, shape, scale=1.0 / shape)
    out = ["+ ----", "Synthetic data file for the generation of sub-grid cloud.",
           "Applicable for spectral files: sp_sw_synth, sp_lw_synth", "- ----", "*DATA", "",
           "tot_subcol_gen", "%5d" % n_gen,
           "subcol_need_single", "%5d" % n_req,
           "subcol_need_optimal", "%5d" % n_req,
           "ipph", "%5d" % 1, "ioverlap", "%5d" % 3, "",
           "lw_subcol_reorder_single"]
    for i in range(0, n_req, 10):
        out.append("".join("%6d" % x for x in reorder[i:i + 10]))
    out.append("")
    out.append("lw_subcol_reorder_optimal")
    for i in range(0, n_req, 10):
        out.append("".join("%6d" % x for x in reorder[i:i + 10]))
    out += ["", "sw_subcol_k"]
    for b, k in sorted(sw_k2):
        out.append("%4d%4d%4d" % (b, k, 2))
    out.append("%4d%4d%4d" % (-99, 0, 0))
    out += ["", "lw_subcol_k"]
    for b, k in sorted(lw_k2):
        out.append("%4d%4d%4d" % (b, k, 2))
    out.append("%4d%4d%4d" % (-99, 0, 0))
    out += ["", "n1", "%6d" % n1, "n2", "%6d" % n2, "xcw"]
    flat = xcw.flatten(order="F")
    for i in range(0, flat.size, 5):
        out.append("".join("%16.8E" % x for x in flat[i:i + 5]))
    out.append("*END")
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")


# Precision-stress spectral files (sp_sw_stress, sp_lw_stress): the band
# structure of the synthetic files (same k-terms per band, so the same McICA
# data apply), one well-mixed major absorber (CO2, whose mixing ratio is the
# call's constant co2_mix_ratio) with no pressure or temperature scaling
# (a = b = 0), no minor gases:
#  * SW: Rayleigh coefficients 10^U(-4, 0) m2/kg per band, so layer optical
#    depths over a 60-layer column (layer masses 0.02..1800 kg m-2) span
#    1e-6..1e3; k-terms chosen so that k co2 / k_rayleigh = 10^U(-8, -1):
#    the clear-sky single-scattering albedo of each g-point is
#    1 - 10^U(-8, -1);
#  * LW: k-terms 10^U(-6, -2) m2/kg (gas optical depths down to 1e-11 at the
#    top, into the thin-layer branch of the infra-red source);
#  * both: Pade-2 droplets with the dev file's extinction fit, a co-albedo
#    10^U(-8, -1) per band and an asymmetry rising linearly from 0 at
#    re = 5 um to 0.95 at re = 20 um, so cloudy McICA sub-columns cover
#    omega -> 1 with asymmetry U(0, 0.95).
STRESS_CO2 = 6.0e-4


def stress_block5(rng, bands, ray):
    out = ["*BLOCK: TYPE =    5: SUBTYPE =    0: VERSION =    1",
           "Exponential sum fiting coefficients: (exponents: m2/kg)",
           "Band        Gas, Number of k-terms, Scaling type and scaling function,",
           "             followed by reference pressure and temperature,",
           "                   k-terms, weights and scaling parameters."]
    for ib, (_, n, _) in enumerate(bands):
        out.append("%5d       %5d       %5d       %5d       %5d" % (ib + 1, 1, n, 2, 1))
        out.append("      %s      %s" % (e16(1.0e5), e16(250.0)))
        if ray is not None:
            k = ray[ib] * 10.0 ** rng.uniform(-8.0, -1.0, n) / STRESS_CO2
        else:
            k = 10.0 ** rng.uniform(-6.0, -2.0, n)
        k.sort()
        w = rng.uniform(0.05, 1.0, n)
        w /= w.sum()
        w = np.array([float("%16.9E" % x) for x in w])
        w[-1] = 1.0 - w[:-1].sum()
        for i in range(n):
            out.append("   %s   %s%3d%s   %s" % (e16(k[i]), e16(w[i]), 0, e16(0.0), e16(0.0)))
    out.append("*END")
    return out


def stress_block10(rng, dev_lines, n_band):
    """Pade-2 droplets (type 5): the dev extinction fit, co-albedo
    10^U(-8, -1) per band, asymmetry linear in re (0 at 5 um, 0.95 at 20 um)."""
    head = dev_lines[:6]
    out = list(head)
    vals = []
    i = 6
    while not dev_lines[i].startswith("*END"):
        if dev_lines[i].startswith("Band"):
            i += 1
            continue
        vals += [float(x) for x in dev_lines[i].split()]
        i += 1
    slope = 0.95 / 15.0e-6
    for b in range(n_band):
        ext = vals[16 * b:16 * b + 6]
        p = ext + [10.0 ** rng.uniform(-8.0, -1.0), 0.0, 0.0, 0.0, 0.0,
                   -5.0e-6 * slope, slope, 0.0, 0.0, 0.0]
        out.append("Band = %5d     Fitting parameters:" % (b + 1))
        for j in range(0, 16, 4):
            out.append("".join("    %12.5E" % x for x in p[j:j + 4]))
    out.append("*END")
    return out


def write_stress_spectrum(path, dev, bands, rng, solar):
    n_band = len(bands)
    lines = block0(n_band, [2])
    lines += read_block(dev, 1)
    ray = None
    if solar:
        lines += read_block(dev, 2)
        ray = 10.0 ** rng.uniform(-4.0, 0.0, n_band)
        b3 = read_block(dev, 3)
        lines += b3[:4] + ["%5d        %s" % (b + 1, "%15.9E" % ray[b]) for b in range(n_band)] + ["*END"]
        ray = np.array([float("%15.9E" % x) for x in ray])
    lines += block4([(2, n, []) for _, n, _ in bands], {2: 1})
    lines += stress_block5(rng, bands, ray)
    lines += block8(n_band)
    if not solar:
        lines += block6_from_dev(read_block(dev, 6))
    lines += stress_block10(rng, read_block(dev, 10), n_band)
    lines += read_block(dev, 12)
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    a = ap.parse_args()
    dev = os.path.join(a.ref, "data", "spectra", "dev")
    os.makedirs(OUT, exist_ok=True)
    rng = np.random.default_rng(20261202)
    write_spectrum(os.path.join(OUT, "sp_lw_synth"), os.path.join(dev, "sp_lw_9_jm3"),
                   LW_BANDS, rng, solar=False)
    write_spectrum(os.path.join(OUT, "sp_sw_synth"), os.path.join(dev, "sp_sw_6_jm2"),
                   SW_BANDS, rng, solar=True)
    write_mcica(os.path.join(OUT, "mcica_synth"), rng, LW_BANDS, SW_BANDS)
    rs = np.random.default_rng(20261203)
    write_stress_spectrum(os.path.join(OUT, "sp_lw_stress"), os.path.join(dev, "sp_lw_9_jm3"),
                          LW_BANDS, rs, solar=False)
    write_stress_spectrum(os.path.join(OUT, "sp_sw_stress"), os.path.join(dev, "sp_sw_6_jm2"),
                          SW_BANDS, rs, solar=True)
    print("wrote", sorted(os.listdir(OUT)))


if __name__ == "__main__":
    main()
