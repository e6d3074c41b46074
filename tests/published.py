"""Figures the reference publishes for its validation set, as printed
(README.md:133-200, docs/report/index.html sections 1 and 3)."""

# Poiseuille relative L2 error: (scheme, q, rows, published)
POISEUILLE = [
    ("bgk", 0.5, 8, 1.112e-2), ("bgk", 0.5, 16, 2.781e-3), ("bgk", 0.5, 32, 6.954e-4),
    ("trt14", 0.5, 8, 7.131e-3), ("trt14", 0.5, 16, 1.783e-3),
    ("trt", 0.25, 8, 3.398e-2), ("trt", 0.25, 16, 7.818e-3),
    ("trt", 0.8, 8, 1.499e-2), ("trt", 0.8, 16, 4.098e-3),
]

# Schaefer-Turek 2D-1 (Re 20): (cells per D, lattice U, c_D, c_L, L_a in m, dP in Pa), as printed
ST_2D1 = [
    (10, 0.04, "5.6132", "0.01050", "0.0802", "0.1133"),
    (20, 0.04, "5.5891", "0.01052", "0.0834", "0.1159"),
    (40, 0.04, "5.5789", "0.01079", "0.0838", "0.1170"),
    (80, 0.04, "5.5774", "0.01079", "0.0840", "0.1173"),
    (20, 0.02, "5.5928", "0.01037", "0.0838", "0.1162"),
]

# Schaefer-Turek 2D-2 (Re 100): (cells per D, lattice U, c_D max, c_D min, c_L max, c_L min, St),
# as printed
ST_2D2 = [
    (20, 0.05, "3.3244", "3.2719", "0.9807", "-1.0175", "0.2988"),
    (40, 0.05, "3.2423", "3.1919", "0.9830", "-1.0181", "0.3006"),
    (80, 0.05, "3.2354", "3.1840", "0.9947", "-1.0295", "0.3007"),
    (20, 0.025, "3.3268", "3.2633", "0.9797", "-1.0167", "0.2995"),
    (40, 0.025, "3.2431", "3.1790", "0.9792", "-1.0144", "0.3013"),
]


def printed(value: float, published: str) -> bool:
    """`value` rounds to the printed figure `published` (a decimal string) at its number of
    decimals."""
    decimals = len(published.split(".")[1])
    return f"{value:.{decimals}f}" == published or abs(value - float(published)) <= 0.5 * 10.0 ** (-decimals)

# Lid-driven cavity (README.md:147-155): (Re, n, max dev from Ghia u, v, u_min, v_max, v_min),
# as printed
CAVITY = [
    (100, 65, "0.0048", "0.0079", "-0.21354", "0.17882", "-0.25245"),
    (100, 129, "0.0050", "0.0084", "-0.21384", "0.17918", "-0.25301"),
    (1000, 65, "0.0167", "0.0087", "-0.37995", "0.36766", "-0.51217"),
    (1000, 129, "0.0045", "0.0115", "-0.38676", "0.37501", "-0.52401"),
    (1000, 257, "0.0061", "0.0158", "-0.38798", "0.37638", "-0.52600"),
]


def ghia(re: int):
    """Ghia's centreline tables from the committed fixture: (u rows, v rows) of (coordinate,
    value), sorted by coordinate."""
    import csv
    import os
    col = {100: 3, 1000: 4}[re]
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ghia1982.csv")
    with open(path) as fh:
        rows = list(csv.reader(line for line in fh if not line.startswith("#")))[1:]
    u = sorted((float(r[2]), float(r[col])) for r in rows if r[0] == "u")
    v = sorted((float(r[2]), float(r[col])) for r in rows if r[0] == "v")
    return u, v

# Corner study: (set, D, Re, U, section, C_D, C_L', St), as printed in the generated report
# (docs/report/index.html, section 4).  README.md:205-229 prints the same tables but three
# figures one unit higher in the last digit (main chamfered C_L' 0.366, main double recess C_D
# 1.352, low-Mach square C_L' 0.385); the report is the one generated from the result files.
CORNERS = [
    ("main", 60, 200.0, 0.1, "square", "1.483", "0.410", "0.1594"),
    ("main", 60, 200.0, 0.1, "chamfered", "1.348", "0.365", "0.1817"),
    ("main", 60, 200.0, 0.1, "rounded", "1.370", "0.354", "0.1778"),
    ("main", 60, 200.0, 0.1, "recessed", "1.364", "0.382", "0.1817"),
    ("main", 60, 200.0, 0.1, "double-recessed", "1.351", "0.375", "0.1825"),
    ("coarse", 40, 200.0, 0.1, "square", "1.465", "0.390", "0.1642"),
    ("coarse", 40, 200.0, 0.1, "chamfered", "1.350", "0.365", "0.1816"),
    ("coarse", 40, 200.0, 0.1, "rounded", "1.374", "0.354", "0.1779"),
    ("coarse", 40, 200.0, 0.1, "recessed", "1.365", "0.385", "0.1820"),
    ("coarse", 40, 200.0, 0.1, "double-recessed", "1.353", "0.377", "0.1827"),
    ("lowmach", 40, 200.0, 0.05, "square", "1.451", "0.384", "0.1648"),
    ("lowmach", 40, 200.0, 0.05, "double-recessed", "1.340", "0.367", "0.1829"),
    ("re100", 40, 100.0, 0.1, "square", "1.500", "0.183", "0.1491"),
    ("re100", 40, 100.0, 0.1, "chamfered", "1.429", "0.182", "0.1554"),
    ("re100", 40, 100.0, 0.1, "rounded", "1.448", "0.178", "0.1534"),
    ("re100", 40, 100.0, 0.1, "recessed", "1.432", "0.188", "0.1563"),
    ("re100", 40, 100.0, 0.1, "double-recessed", "1.425", "0.184", "0.1564"),
]
