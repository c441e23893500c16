"""Test-side statement of the Harris-sheet cases (HarrisSheetCase,
src/initialization/cases/harris_sheet.cpp:19-257 of the reference): the seed,
the curl parameters of currentFromAmpere, the per-species profile and the
sheet field, for building the oracle's runs of Harris decks. The relation
between the field and its curl is checked by test_oracle (Ampere's law on
the loaded particles against a finite-difference curl of magnetic())."""
import math

import numpy as np

import oracle as O

PI = 3.1415926535897932384626433832795
TWO_PI = 6.283185307179586476925286766559


def seed(case, lx, ly, delta, pert_gem=0.0, pert_hump=0.0, deltax_hump=8.0, deltay_hump=4.0, kx_hump=-1.0,
         ky_hump=-1.0):
    s = dict(mode_amplitude=0.0, mode_kx=0.0, mode_ky=0.0, amplitude_x=0.0, amplitude_y=0.0, width_x=0.0,
             width_y=0.0, period_x=0.0, period_y=0.0, slope_period_x=0.0, slope_period_y=0.0)
    if case == "GEMHarris":
        s.update(mode_amplitude=pert_gem, mode_kx=TWO_PI * 1 / lx, mode_ky=TWO_PI * 1 / ly)
        if pert_hump != 0.0:
            kx = TWO_PI / lx if kx_hump < 0.0 else kx_hump
            ky = .5 * TWO_PI / ly if ky_hump < 0.0 else ky_hump
            wx, wy = deltax_hump * delta, deltay_hump * delta
            s.update(width_x=wx, width_y=wy, amplitude_x=pert_hump / (ky * wy), amplitude_y=pert_hump / (ky * wx),
                     period_x=PI / (kx * delta), period_y=PI / (ky * delta), slope_period_x=PI / (kx * wx),
                     slope_period_y=PI / (ky * wy))
    elif case in ("GEM", "DoubleGEM"):
        s.update(amplitude_x=0.4, amplitude_y=0.4, width_x=delta, width_y=delta, period_x=10.0, period_y=10.0,
                 slope_period_x=10.0, slope_period_y=10.0)
    return s


def has_hump(s):
    return s["amplitude_x"] != 0.0 or s["amplitude_y"] != 0.0


def magnetic(case, s, b0x, delta, lx, ly, x, y):
    """HarrisSheetCase::magnetic (harris_sheet.cpp:229-255), without B0."""
    double = case in ("DoubleGEM", "DoubleGEMnoPert")
    if double:
        bx = b0x * (-1.0 + math.tanh((y - 0.25 * ly) / delta) - math.tanh((y - 0.75 * ly) / delta))
    else:
        bx = b0x * math.tanh((y - 0.5 * ly) / delta)
    by = 0.0
    if s["mode_amplitude"] != 0.0:
        xm, ym = x - 0.5 * lx, y - 0.5 * ly
        bx += -s["mode_amplitude"] * b0x * (lx / ly) * math.cos(s["mode_kx"] * xm) * math.sin(s["mode_ky"] * ym)
        by += s["mode_amplitude"] * b0x * math.sin(s["mode_kx"] * xm) * math.cos(s["mode_ky"] * ym)
    if has_hump(s):
        centers = [(0.25, 0.25, 1.0), (0.75, 0.75, -1.0)] if double else [(0.5, 0.5, 1.0)]
        for fx, fy, sign in centers:
            xp, yp = x - fx * lx, y - fy * ly
            env = math.exp(-(xp / s["width_x"]) * (xp / s["width_x"]) - (yp / s["width_y"]) * (yp / s["width_y"]))
            cx = math.cos(PI * xp / s["period_x"] / delta)
            sx = math.sin(PI * xp / s["period_x"] / delta)
            cy = math.cos(PI * yp / s["period_y"] / delta)
            sy = math.sin(PI * yp / s["period_y"] / delta)
            bx += sign * b0x * s["amplitude_x"] * env * (-cx * cy * 2.0 * yp / s["width_y"] -
                                                         cx * sy * PI / s["slope_period_y"])
            by += sign * b0x * s["amplitude_y"] * env * (cx * cy * 2.0 * xp / s["width_x"] +
                                                         sx * cy * PI / s["slope_period_x"])
    return bx, by


def profiles(case, s, b0x, delta, lx, ly, species, current_from_ampere=False, c=1.0):
    """Per species an O.Harris structure (HarrisSheetCase constructor and
    admit_ampere_current). species: dicts with qom, rho, drift."""
    double = case in ("DoubleGEM", "DoubleGEMnoPert")
    seen = {True: False, False: False}
    out = []
    for sp in species:
        if case == "plasmoidChain":
            electron = sp["qom"] < 0.0
            density = not seen[electron]
            seen[electron] = True
        else:
            density = any(abs(d) > 1.0e-10 for d in sp["drift"])
        out.append(O.harris(density=int(density), is_double=int(double), ly=ly, delta=delta))
    if case == "GEMHarris" and current_from_ampere:
        drift_sum = sum((1.0 if sp["qom"] > 0.0 else -1.0) * sp["drift"][2] for sp in species)
        curl = dict(center_x=0.5 * lx, center_y=0.5 * ly, curl_delta=delta, sheet=-b0x / delta,
                    mode=(s["mode_amplitude"] * b0x) * (s["mode_kx"] + (lx / ly) * s["mode_ky"]),
                    mode_kx=s["mode_kx"], mode_ky=s["mode_ky"])
        if has_hump(s):
            curl.update(hump=(b0x * s["amplitude_x"]) * s["width_y"], hump_kx=PI / (s["period_x"] * delta),
                        hump_ky=PI / (s["period_y"] * delta), hump_inverse_wx2=1.0 / (s["width_x"] * s["width_x"]),
                        hump_inverse_wy2=1.0 / (s["width_y"] * s["width_y"]))
        for h, sp in zip(out, species):
            if sp["drift"][2] == 0.0:
                continue
            h.ampere_drift = (sp["drift"][2] * c) / (drift_sum * sp["rho"])
            for k, v in curl.items():
                setattr(h, k, v)
    return out


def cell_field(case, s, b0x, delta, cells, lengths, b0=(0.0, 0.0, 0.0)):
    """B at the cell centres (x fastest), B0 added."""
    nx, ny, nz = cells
    dx, dy = lengths[0] / nx, lengths[1] / ny
    b = np.zeros((3, nz, ny, nx))
    for j in range(ny):
        for i in range(nx):
            bx, by = magnetic(case, s, b0x, delta, lengths[0], lengths[1], (i + 0.5) * dx, (j + 0.5) * dy)
            b[0, :, j, i] = bx + b0[0]
            b[1, :, j, i] = by + b0[1]
            b[2, :, j, i] = 0.0 + b0[2]
    return b.ravel()
