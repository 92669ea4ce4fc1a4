"""Generate let_like.json: the LET-like direct-geometry chopper spectrometer
used as the headline benchmark workload.

- Source: MCont, T1 = T2 = T3 = 40 K, I1 = 1, I2 = I3 = 0, with the 1/E
  epithermal tail joined at 2 k_B T (6.89 meV, inside the window); E in
  [1.1, 9.0] meV; 0.04 x 0.09 m face
  aimed at the 0.04 x 0.09 m guide entrance at 1.7 m.
- Guide: 20 straight supermirror segments between z = 1.7 m and 26.7 m, 2 mm
  gaps between segments, width tapering linearly 0.04 -> 0.02 m, height
  0.09 m, m = 2 (R0 0.99, Qc 0.0219 1/A, alpha 6.07, W 0.003).
- Choppers: 5 disks (radius 0.29 m) in the gaps at z = 7.83, 8.40, 11.75,
  15.66, 23.50 m with 1, 1, 2, 1, 2 openings at 300, 150, 20, 100, 300 Hz
  (freq = 2 pi f rad/s), opening widths 0.04-0.06 m, phased to open at
  t = z / 841.3 m/s (Ei = 3.7 meV for a source emitting at t = 0).
- Sample: incoherent cylinder r = 5 mm, h = 20 mm at z = 28.0 m,
  Sigma_s = 0.36 /cm, Sigma_a = 0.26 lambda/1.8 /cm.
- Detector: cylinder of radius 3.5 m around the sample; 768 tubes over
  2theta in [-40, 140] deg x 128 pixels over y in [-2, 2] m x 1000 TOF bins
  over [10, 60] ms (98,304 pixels, 98.3 M bins).
"""
import json
import os

V_FOC = 841.3
CHOPPERS = [  # z, n_slits, f (Hz), slit width (m)
    (7.83, 1, 300.0, 0.04),
    (8.40, 1, 150.0, 0.05),
    (11.75, 2, 20.0, 0.06),
    (15.66, 1, 100.0, 0.05),
    (23.50, 2, 300.0, 0.04),
]
Z0, Z1 = 1.7, 26.7
N_SEG = 20
GAP = 0.002


def width(z):
    return 0.04 - 0.02 * (z - Z0) / (Z1 - Z0)


def boundaries():
    b = [Z0, Z1] + [c[0] for c in CHOPPERS]
    b = sorted(b)
    while len(b) < N_SEG + 1:
        k = max(range(len(b) - 1), key=lambda i: b[i + 1] - b[i])
        b.insert(k + 1, 0.5 * (b[k] + b[k + 1]))
    return b


def build():
    comps = {}
    comps["linear"] = True
    comps["source"] = {
        "source": True,
        "position": [0.0, 0.0, 0.0],
        "moderator_kernel": {
            "name": "MCont", "mod_dim": [0.04, 0.09], "target_dim": [0.04, 0.09],
            "target_dist": 1.7, "E_min": 1.1, "E_max": 9.0,
            "T1": 40.0, "I1": 1.0, "T2": 40.0, "I2": 0.0, "T3": 40.0, "I3": 0.0,
            "tail_join": 2.0,
        },
    }
    b = boundaries()
    chop_at = {round(c[0], 6): c for c in CHOPPERS}
    for i in range(N_SEG):
        za = b[i] + (GAP / 2 if i > 0 else 0.0)
        zb = b[i + 1] - (GAP / 2 if i < N_SEG - 1 else 0.0)
        w1, w2 = width(za), width(zb)
        comps[f"guide{i + 1:02d}"] = {
            "position": [0.0, 0.0, round(za, 6)],
            "geom_kernel": {"name": "GPlane", "width": round(w1, 6), "height": 0.09},
            "scat_kernel": {"name": "SGuide", "w1": round(w1, 6), "h1": 0.09, "w2": round(w2, 6),
                            "h2": 0.09, "l": round(zb - za, 6), "R0": 0.99, "Qc": 0.0219,
                            "alpha": 6.07, "m": 2, "W": 0.003},
        }
        c = chop_at.get(round(b[i + 1], 6))
        if c is not None and i < N_SEG - 1:
            z, ns, f, sw = c
            comps[f"chopper_{z:.2f}"] = {
                "position": [0.0, 0.0, z],
                "geom_kernel": {"name": "GPlane", "width": 0.6, "height": 0.6},
                "scat_kernel": {"name": "SChopper", "radius": 0.29, "freq": 2 * 3.141592653589793 * f,
                                "n_slits": ns, "jitter": 0.0, "slit_width": sw,
                                "phase": z / V_FOC},
            }
    comps["sample"] = {
        "position": [0.0, 0.0, 28.0],
        "geom_kernel": {"name": "GCylinderExt", "radius": 0.005, "height": 0.02},
        "scat_kernel": {"name": "SIncoherent", "sigma_s": 0.36, "sigma_a": 0.26, "lambda_ref": 1.8},
    }
    comps["bank"] = {
        "position": [0.0, 0.0, 28.0],
        "geom_kernel": {"name": "GBanana", "radius": 3.5, "height": 4.0, "mintheta": -40.0,
                        "maxtheta": 140.0},
        "scat_kernel": {"name": "SDetectorTubes", "theta_range": [-40.0, 140.0], "y_range": [-2.0, 2.0],
                        "tof_range": [10000.0, 60000.0], "n_tubes": 768, "n_pixels": 128, "n_tof": 1000},
    }
    return {"let_like": comps}


if __name__ == "__main__":
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "let_like.json")
    with open(out, "w") as f:
        json.dump(build(), f, indent=1)
    print("wrote", out)
