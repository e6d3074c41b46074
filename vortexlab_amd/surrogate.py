"""Learned surrogate for the towing-tank sweep: body mask + Re -> (mean Cd, Cl', St).

The labels come from the batched solver (vl_batch); the network is trained and evaluated by
the HIP kernels behind vl_nn_* (vortexlab_amd/csrc/nn.hip).  A 4-member ensemble answers a
query when its members agree on the drag (relative spread of the mean Cd at most a threshold
calibrated on shapes with known answers so that the accepted answers meet a stated 90th-percentile
error), otherwise the query falls back to the solver, so the ensemble's spread also decides when its
own answer is trusted.

Usage (one GPU):  python -m vortexlab_amd.surrogate --train 8192 --test 1024 --out report.json
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import time
from dataclasses import dataclass

import numpy as np

from . import api, workloads

IMG = 64
N_OUT = 3
OUT_NAMES = ("cd_mean", "cl_rms", "st")
_F = C.POINTER(C.c_float)
_I = C.POINTER(C.c_int32)


def _bind(lib):
    sig = {
        "vl_nn_create": (C.c_int, [C.c_int, C.c_uint64, C.POINTER(C.c_void_p)]),
        "vl_nn_destroy": (None, [C.c_void_p]),
        "vl_nn_param_count": (C.c_int64, []),
        "vl_nn_get_params": (C.c_int, [C.c_void_p, _F]),
        "vl_nn_set_params": (C.c_int, [C.c_void_p, _F]),
        "vl_nn_forward": (C.c_int, [C.c_void_p, _F, _F, C.c_int, _F]),
        "vl_nn_gradients": (C.c_int, [C.c_void_p, _F, _F, _F, C.c_int, _F, _F]),
        "vl_nn_set_dataset": (C.c_int, [C.c_void_p, _F, _F, _F, C.c_int64]),
        "vl_nn_train_epoch": (C.c_int, [C.c_void_p, _I, C.c_int64, C.c_int, C.c_float, _F]),
        "vl_nn_predict_dataset": (C.c_int, [C.c_void_p, _I, C.c_int64, _F]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype, fn.argtypes = res, args


def _fp(a):
    return a.ctypes.data_as(_F)


class Net:
    """One network instance on the GPU (parameters, Adam state, activations, dataset)."""

    def __init__(self, be: api.Backend, seed: int, max_batch: int = 128):
        self.be = be
        _bind(be.lib)
        h = C.c_void_p()
        be.check(be.lib.vl_nn_create(max_batch, seed, C.byref(h)))
        self.h = h.value
        self.max_batch = max_batch
        self.n_params = int(be.lib.vl_nn_param_count())

    def params(self) -> np.ndarray:
        p = np.zeros(self.n_params, dtype=np.float32)
        self.be.check(self.be.lib.vl_nn_get_params(self.h, _fp(p)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        self.be.check(self.be.lib.vl_nn_set_params(self.h, _fp(p)))

    def forward(self, masks, feat) -> np.ndarray:
        masks = np.ascontiguousarray(masks, dtype=np.float32).reshape(-1, IMG * IMG)
        feat = np.ascontiguousarray(feat, dtype=np.float32).reshape(-1)
        out = np.zeros((len(feat), N_OUT), dtype=np.float32)
        for b0 in range(0, len(feat), self.max_batch):
            n = min(self.max_batch, len(feat) - b0)
            o = np.zeros((n, N_OUT), dtype=np.float32)
            self.be.check(self.be.lib.vl_nn_forward(self.h, _fp(masks[b0:b0 + n]), _fp(feat[b0:b0 + n]), n, _fp(o)))
            out[b0:b0 + n] = o
        return out

    def gradients(self, masks, feat, targets):
        masks = np.ascontiguousarray(masks, dtype=np.float32).reshape(-1, IMG * IMG)
        feat = np.ascontiguousarray(feat, dtype=np.float32).reshape(-1)
        targets = np.ascontiguousarray(targets, dtype=np.float32).reshape(-1, N_OUT)
        loss = C.c_float()
        g = np.zeros(self.n_params, dtype=np.float32)
        self.be.check(self.be.lib.vl_nn_gradients(self.h, _fp(masks), _fp(feat), _fp(targets), len(feat),
                                                  C.byref(loss), _fp(g)))
        return float(loss.value), g

    def set_dataset(self, masks, feat, targets):
        self._ds = [np.ascontiguousarray(masks, dtype=np.float32).reshape(-1, IMG * IMG),
                    np.ascontiguousarray(feat, dtype=np.float32).reshape(-1),
                    np.ascontiguousarray(targets, dtype=np.float32).reshape(-1, N_OUT)]
        self.be.check(self.be.lib.vl_nn_set_dataset(self.h, *[_fp(a) for a in self._ds], len(self._ds[1])))

    def train_epoch(self, order, batch: int, lr: float) -> float:
        order = np.ascontiguousarray(order, dtype=np.int32)
        loss = C.c_float()
        self.be.check(self.be.lib.vl_nn_train_epoch(self.h, order.ctypes.data_as(_I), len(order), batch, lr,
                                                    C.byref(loss)))
        return float(loss.value)

    def predict_dataset(self, idx) -> np.ndarray:
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        out = np.zeros((len(idx), N_OUT), dtype=np.float32)
        self.be.check(self.be.lib.vl_nn_predict_dataset(self.h, idx.ctypes.data_as(_I), len(idx), _fp(out)))
        return out

    def close(self):
        if self.h:
            self.be.lib.vl_nn_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------------------
# masks
# ---------------------------------------------------------------------------------------
def rasterize(outline: np.ndarray, cx: float, cy: float, d: float, half: float = 0.6) -> np.ndarray:
    """64 x 64 mask of a polygon over the square of side 2 * half * d centred on (cx, cy) (the
    star bodies reach at most d/2 from the centre): a pixel is 1 when its centre is inside by
    the even-odd rule (the solver's polygon predicate)."""
    side = 2.0 * half * d
    c = (np.arange(IMG) + 0.5) / IMG * side - 0.5 * side
    px, py = np.meshgrid(cx + c, cy + c)
    inside = np.zeros(px.shape, dtype=bool)
    x, y = outline[:, 0], outline[:, 1]
    j = len(x) - 1
    for i in range(len(x)):
        xi, yi, xj, yj = x[i], y[i], x[j], y[j]
        cross = (yi > py) != (yj > py)
        with np.errstate(divide="ignore", invalid="ignore"):
            xint = (xj - xi) * (py - yi) / (yj - yi) + xi
        inside ^= cross & (px < xint)
        j = i
    return inside.astype(np.float32)


# ---------------------------------------------------------------------------------------
# dataset from the batched solver
# ---------------------------------------------------------------------------------------
@dataclass
class Dataset:
    masks: np.ndarray   # [n, 64, 64]
    re: np.ndarray      # [n]
    y: np.ndarray       # [n, 3] (cd_mean, cl_rms, st)
    ok: np.ndarray      # [n] solver finished (did not diverge)
    solve_seconds: float
    lattice_updates: int


def sample_cases(seed: int, n: int, d: float = 24.0, re=(40.0, 250.0), units: float = 30.0, u: float = 0.1,
                 upstream: float = 4.0, downstream: float = 12.0, height: float = 6.0):
    """n towing-tank cases with star-polygon bodies (workloads.star_polygon), Re uniform in
    `re`, TRT, in a fixed channel of (upstream + downstream) x height diameters.  With 24 cells
    per diameter and stream speed 0.1 (Mach 0.17), tau >= 0.529 at Re 250."""
    rng = np.random.default_rng(seed)
    nx = int(round((upstream + downstream) * d)) + 2
    ny = int(round(height * d)) + 2
    cases, masks, res = [], [], []
    for _ in range(n):
        r = float(rng.uniform(*re))
        sub = np.random.default_rng(rng.integers(1 << 62))
        holder = {}

        def outline_fn(cx, cy, sub=sub, holder=holder):
            holder["o"] = workloads.star_polygon(sub, d, cx, cy)
            holder["c"] = (cx, cy)
            return holder["o"]

        c = workloads.tank_case(nx, ny, d, r, u, api.TRT, outline_fn, units=units, ramp_units=3.0, kick_units=1.0)
        cx, cy = holder["c"]
        cases.append(c)
        masks.append(rasterize(holder["o"], cx, cy, d))
        res.append(r)
    return cases, np.stack(masks), np.asarray(res)


def solve(be: api.Backend, cases, discard_units: float = 10.0, precision: int = api.F32, cd_limit: float = 50.0):
    """Runs the cases in one batch and returns their (cd_mean, cl_rms, st) and status.  A run
    whose drag leaves [-cd_limit, cd_limit] is counted as diverged, like a non-finite one: a
    blow-up that has not reached infinity by the last step is still a blow-up."""
    b = api.Batch(be, precision)
    for c in cases:
        b.add(c.spec, c.steps, c.d)
    b.build()
    t0 = time.perf_counter()
    b.run()
    secs = time.perf_counter() - t0
    lu = int(b.info().lattice_updates)
    y = np.zeros((len(cases), N_OUT))
    ok = np.zeros(len(cases), dtype=bool)
    for k, c in enumerate(cases):
        coef, done, st = b.result(k)
        if st != 0 or not np.all(np.isfinite(coef)) or np.max(np.abs(coef[:, 0])) > cd_limit:
            continue
        s = api.series_stats(be, coef[:, 0], coef[:, 1], c.spec.u / c.d, discard_units)
        y[k] = (s.cd_mean, s.cl_rms, s.st)
        ok[k] = True
    b.close()
    return y, ok, secs, lu


def save_datasets(path: str, **sets: Dataset) -> None:
    """Writes solver-labelled datasets to one .npz file (arrays only, no pickled objects)."""
    arrs = {}
    for name, ds in sets.items():
        arrs.update({f"{name}_masks": ds.masks, f"{name}_re": ds.re, f"{name}_y": ds.y, f"{name}_ok": ds.ok,
                     f"{name}_meta": np.array([ds.solve_seconds, ds.lattice_updates], dtype=np.float64)})
    np.savez(path, **arrs)


def load_datasets(path: str, *names: str):
    """The datasets `names` from a file written by save_datasets."""
    with np.load(path) as f:
        out = []
        for name in names:
            meta = f[f"{name}_meta"]
            out.append(Dataset(f[f"{name}_masks"], f[f"{name}_re"], f[f"{name}_y"], f[f"{name}_ok"], float(meta[0]),
                               int(meta[1])))
    return out


def make_dataset(be, seed, n, **kw) -> Dataset:
    cases, masks, re = sample_cases(seed, n, **kw)
    y, ok, secs, lu = solve(be, cases)
    return Dataset(masks, re, y, ok, secs, lu)


# ---------------------------------------------------------------------------------------
# ensemble
# ---------------------------------------------------------------------------------------
class Ensemble:
    """Members trained on the same data from different seeds and sample orders."""

    def __init__(self, be, members: int = 4, seed: int = 1, max_batch: int = 128):
        self.nets = [Net(be, seed + 7919 * k, max_batch) for k in range(members)]
        self.seed = seed

    def fit(self, masks, re, y, epochs: int = 40, batch: int = 64, lr: float = 1e-3, log=None):
        self.feat_mu, self.feat_sd = float(np.mean(np.log(re))), float(np.std(np.log(re)) + 1e-12)
        self.y_mu, self.y_sd = y.mean(0), y.std(0) + 1e-9
        feat = (np.log(re) - self.feat_mu) / self.feat_sd
        yt = (y - self.y_mu) / self.y_sd
        losses = []
        for k, net in enumerate(self.nets):
            net.set_dataset(masks, feat, yt)
            rng = np.random.default_rng(self.seed * 1000 + k)
            for ep in range(epochs):
                lr_ep = lr * 0.5 * (1.0 + math.cos(math.pi * ep / epochs))  # cosine decay
                loss = net.train_epoch(rng.permutation(len(feat)), batch, lr_ep)
            losses.append(loss)
            if log:
                log(f"[surrogate] member {k}: final epoch loss {loss:.4f}")
        return losses

    def predict(self, masks, re):
        feat = (np.log(re) - self.feat_mu) / self.feat_sd
        outs = np.stack([net.forward(masks, feat) for net in self.nets]) * self.y_sd + self.y_mu
        return outs.mean(0), outs.std(0)

    def close(self):
        for n in self.nets:
            n.close()


def spread_of(mean, std, y_scale, output: int = 0):
    """The ensemble's relative spread on one output: standard deviation over the members / |mean|
    (or / a tenth of the output's dataset scale when the mean is near zero)."""
    return std[:, output] / np.maximum(np.abs(mean[:, output]), 0.1 * y_scale[output])


def accept(mean, std, y_scale, threshold: float = 0.02, outputs=(0,)):
    """A prediction is accepted when the ensemble's relative spread is at most `threshold` on
    each of `outputs` (default: the drag, the primary quantity)."""
    ok = np.stack([spread_of(mean, std, y_scale, o) <= threshold for o in outputs], axis=1)
    return np.all(ok, axis=1)


def calibrate_threshold(spread, rel_err, target_p90: float, quantile: float = 0.9) -> float:
    """The largest spread threshold whose accepted answers (spread <= threshold) have their
    `quantile` relative error at most `target_p90` on a calibration set with known answers; 0 (no
    answer accepted, every query goes to the solver) when not even the surest answer meets it."""
    order = np.argsort(spread, kind="stable")
    s, e = np.asarray(spread)[order], np.asarray(rel_err)[order]
    best = 0.0
    for n in range(1, len(s) + 1):
        if n < len(s) and s[n] == s[n - 1]:
            continue  # a threshold accepts every answer of equal spread
        if np.quantile(e[:n], quantile) <= target_p90:
            best = float(s[n - 1])
    return best


def fallback_report(spread, rel_err, threshold: float) -> dict:
    """What the fallback rule gives on answers with known errors: the accepted fraction, the
    median and 90th-percentile relative error of the accepted answers, and the 90th percentile
    over all queries when the rejected ones are solved (error 0)."""
    acc = spread <= threshold
    combined = np.where(acc, rel_err, 0.0)
    return {"threshold": float(threshold), "accepted_fraction": float(acc.mean()),
            "median_rel_err_accepted": float(np.median(rel_err[acc])) if acc.any() else None,
            "p90_rel_err_accepted": float(np.quantile(rel_err[acc], 0.9)) if acc.any() else None,
            "p90_rel_err_all_queries": float(np.quantile(combined, 0.9)),
            "p90_rel_err_surrogate_alone": float(np.quantile(rel_err, 0.9))}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--train", type=int, default=8192)
    ap.add_argument("--test", type=int, default=1024)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--seed", type=int, default=5)
    ap.add_argument("--members", type=int, default=8, help="ensemble members")
    ap.add_argument("--target-p90", type=float, default=0.05,
                    help="stated error of accepted answers: their 90th-percentile relative Cd error")
    ap.add_argument("--out", default="")
    ap.add_argument("--save-datasets", default="", help="write the solver-labelled sets to this .npz file")
    ap.add_argument("--load-datasets", default="",
                    help="train and evaluate on the sets of a --save-datasets file instead of solving")
    args = ap.parse_args()
    import sys

    import vortexlab_amd as vl

    def log(m):
        print(m, file=sys.stderr, flush=True)

    be = vl.backend()
    t0 = time.perf_counter()
    if args.load_datasets:
        tr, te = load_datasets(args.load_datasets, "train", "test")
        args.train, args.test = len(tr.ok), len(te.ok)
    else:
        tr = make_dataset(be, args.seed, args.train)
        te = make_dataset(be, args.seed + 1, args.test)
    log(f"[surrogate] training set: {tr.ok.sum()} of {args.train} cases solved in {tr.solve_seconds:.1f} s "
        f"({tr.lattice_updates / tr.solve_seconds / 1e9:.1f} GLUPS)")
    log(f"[surrogate] held-out set: {te.ok.sum()} of {args.test} cases solved in {te.solve_seconds:.1f} s")
    if args.save_datasets:
        save_datasets(args.save_datasets, train=tr, test=te)
    ens = Ensemble(be, args.members, seed=args.seed)
    t1 = time.perf_counter()
    ens.fit(tr.masks[tr.ok], tr.re[tr.ok], tr.y[tr.ok], epochs=args.epochs, log=log)
    t_train = time.perf_counter() - t1
    t2 = time.perf_counter()
    mean, std = ens.predict(te.masks[te.ok], te.re[te.ok])
    t_pred = time.perf_counter() - t2
    truth = te.y[te.ok]
    rel = np.abs(mean - truth) / np.maximum(np.abs(truth), 0.1 * ens.y_sd)
    spread = spread_of(mean, std, ens.y_sd)
    # the held-out shapes in two halves: the fallback threshold is calibrated on the first so that
    # its accepted answers meet the stated p90 error, and judged on the second
    half = len(truth) // 2
    cal, ev = slice(0, half), slice(half, len(truth))
    threshold = calibrate_threshold(spread[cal], rel[cal, 0], args.target_p90)
    rep_cal = fallback_report(spread[cal], rel[cal, 0], threshold)
    rep_ev = fallback_report(spread[ev], rel[ev, 0], threshold)
    # how well the spread ranks the answers by their error (Spearman's rank correlation)
    rank_corr = float(np.corrcoef(np.argsort(np.argsort(spread)), np.argsort(np.argsort(rel[:, 0])))[0, 1])
    log(f"[surrogate] threshold {threshold:.4f} for p90 <= {args.target_p90}: calibration accepts "
        f"{rep_cal['accepted_fraction']:.2f}; evaluation half accepts {rep_ev['accepted_fraction']:.2f} with p90 "
        f"{rep_ev['p90_rel_err_accepted']}")
    sweep = {}
    for th in (0.005, 0.01, 0.02, 0.05, 0.1):
        sweep[str(th)] = fallback_report(spread[ev], rel[ev, 0], th)
    baseline = np.abs(ens.y_mu - truth) / np.maximum(np.abs(truth), 0.1 * ens.y_sd)
    solve_per_case = te.solve_seconds / args.test
    pred_per_case = t_pred / max(1, len(truth))
    frac = rep_ev["accepted_fraction"]
    per_query = frac * pred_per_case + (1.0 - frac) * (solve_per_case + pred_per_case)
    report = {
        "train_cases": int(tr.ok.sum()), "test_cases": int(te.ok.sum()), "epochs": args.epochs,
        "train_seconds": round(t_train, 2), "dataset_solve_seconds": round(tr.solve_seconds + te.solve_seconds, 2),
        "rule": f"accept when the ensemble spread of Cd is <= the threshold calibrated on half the held-out "
                f"shapes for a p90 relative Cd error <= {args.target_p90:.0%}; otherwise solve",
        "target_p90_rel_err_cd": args.target_p90,
        "members": args.members, "spread_error_rank_correlation": round(rank_corr, 4),
        "calibration_half": rep_cal, "evaluation_half": rep_ev,
        "accepted_fraction": round(frac, 4),
        "median_rel_error_all": {k: float(np.median(rel[ev, i])) for i, k in enumerate(OUT_NAMES)},
        "p90_rel_error_all": {k: float(np.quantile(rel[ev, i], 0.9)) for i, k in enumerate(OUT_NAMES)},
        "median_rel_error_mean_predictor": {k: float(np.median(baseline[ev, i])) for i, k in enumerate(OUT_NAMES)},
        "solver_seconds_per_case_batched": solve_per_case,
        "surrogate_seconds_per_case": pred_per_case,
        "queries_per_hour_solver_only": 3600.0 / solve_per_case,
        "queries_per_hour_with_surrogate": 3600.0 / per_query,
        "solves_avoided_per_hour": frac * 3600.0 / per_query,
        "threshold_sweep_evaluation_half": sweep,
        "total_seconds": round(time.perf_counter() - t0, 1),
    }
    ens.close()
    print(json.dumps(report), flush=True)
    if args.out:
        with open(args.out, "w") as fh:
            json.dump(report, fh, indent=1)


if __name__ == "__main__":
    main()
