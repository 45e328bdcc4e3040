#!/usr/bin/env python3
"""Benchmark: SNP-pair F-tests/sec of the exhaustive pairwise epistasis scan on one MI355X.

Workload (default c3): N=8000 individuals, M=50000 SNPs on one chromosome,
MAF ~ U(0.05,0.5), no missing calls, phenotype (Gamma(2,1)-2)/sqrt(2) with 10
planted interacting pairs (+0.15 SD in one 3x3 cell), 8df + 4df tests at the
reference's default thresholds 6.5 / 10.5. One step = one full scan of all
M(M-1)/2 = 1,249,975,000 pairs with the genotypes and phenotype resident on
the device: tile list, pair kernels, hit compaction and hit read-back.

Prints one JSON line on stdout; progress goes to stderr. Besides the headline
(MFMA contraction form) it times one scan of the plain popcount form, checks
a block of pairs against the FP64 oracle, and times the FP32 CPU restatement
of the reference kernel on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "SNP-pair F-tests/sec at N=8000,M=50000; MFMA vs popcount form, F relerr≤1e-5"
INT8_PEAK_TOPS = 5000.0  # MI355X dense INT8 MFMA, TOP/s


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(wl, seconds, threads):
    """FP32 CPU restatement of the reference kernel (oracle/epi_refscan32.c) on seeded 128x128 tiles."""
    import oracle as O
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
    rng = np.random.default_rng(12345)
    T = 128
    nt = wl.nsnp // T

    def tiles(k):
        out = []
        while len(out) < k:
            a, b = rng.integers(0, nt, 2)
            if a > b:
                out.append((a * T, b * T))
        return np.array(out, np.int32)

    t0 = time.perf_counter()
    npairs, _ = O.ref32_scan_tiles(wl.words, wl.nid, wl.y[0], tiles(threads), T, wl.test, wl.t1, wl.t2)
    dt = time.perf_counter() - t0
    k = max(threads, int(threads * max(1.0, seconds / max(dt, 1e-3))))
    tl = tiles(k)
    t0 = time.perf_counter()
    npairs, nh = O.ref32_scan_tiles(wl.words, wl.nid, wl.y[0], tl, T, wl.test, wl.t1, wl.t2)
    dt = time.perf_counter() - t0
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": npairs / dt, "unit": "pair-tests/s", "cores": threads, "kind": "port",
            "sample": f"{len(tl)} seeded 128x128 off-diagonal tiles of the workload ({npairs} pairs, {dt:.1f} s)",
            "cpu_model": model, "host_cpus": os.cpu_count()}


# HBM bytes per launch of the dominant kernel, from the committed rocprofv3 summary of the same workload and
# tile height (FETCH_SIZE + WRITE_SIZE, each scaled by its factor on a copy of known size). Both heights were
# profiled with 16 x 2 super-tiles; other super-tile shapes of the same height differ only in L2 reuse, and HBM
# runs at about 1% of its bandwidth either way.
PROFILED = {("c3", "mfma", 64): ("profiles/mfma_fast_kernel_c3.md", "TRAFFIC_BYTES_PER_LAUNCH"),
            ("c3", "mfma", 32): ("profiles/mfma_fast_kernel_c3.md", "TRAFFIC_BYTES_PER_LAUNCH_32")}


def profiled_traffic(workload, form, sched):
    path, key = PROFILED.get((workload, form, sched["fast_rows"]), (None, None))
    if not path or not os.path.exists(os.path.join(ROOT, path)):
        return None
    for line in open(os.path.join(ROOT, path)):
        f = line.split()
        if len(f) == 2 and f[0] == key:
            return float(f[1])
    return None


FP32_PEAK_TFLOPS = 157.3  # MI355X vector FP32


def roofline(form, wl, kernel_ms, traffic=None, source=None):
    """MFMA form: 36 N logical INT8 ops per pair (9 cells x N x (count + sum) MACs) against the dense INT8 peak.
    Plain form: 18 N logical FP32 flops per pair (9 cells x N sum FMAs) against the vector FP32 peak."""
    if form == "mfma":
        ops, peak, bound, note = 36.0 * wl.nid * wl.pairs, INT8_PEAK_TOPS, "mfma", "INT8 ops; 36*N logical ops per pair"
    else:
        ops, peak, bound, note = (18.0 * wl.nid * wl.pairs, FP32_PEAK_TFLOPS, "valu",
                                  "FP32; 18*N logical flops per pair (9 cell-sum FMAs per individual); FAST tiles "
                                  "execute 4 of the 9 and derive the rest from per-SNP marginals")
    achieved = ops / (kernel_ms * 1e-3) / 1e12
    out = {"bound": bound, "achieved": achieved, "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak,
           "traffic": traffic, "traffic_source": source, "ops_per_pair": ops / wl.pairs, "note": note}
    if form != "mfma":
        # what the plain kernel executes: FAST tiles (no missing call) 4 cell FMAs per individual, GENERAL
        # tiles 9 cell FMAs and one of y^2; a workload with missing calls has them in every tile
        exe = 2.0 * wl.nid * (4 if wl.gmiss == 0 else 10)
        out["executed_flops_per_pair"] = exe
        out["executed_achieved"] = exe * wl.pairs / (kernel_ms * 1e-3) / 1e12
        out["executed_frac"] = out["executed_achieved"] / peak
    return out


def accuracy(s, wl, form, phen=0):
    """One 64x64 block of pairs against the FP64 oracle."""
    import oracle as O
    a0, b0 = wl.nsnp // 2 + 17, wl.nsnp // 3 + 5
    counts, sums, F, F2 = s.pair_dump(a0, a0 + 64, b0, b0 + 64, test=wl.test, form=form, phen=phen)
    codes = O.unpack(wl.words[np.r_[a0:a0 + 64, b0:b0 + 64]], wl.nid)
    aa, bb = np.meshgrid(np.arange(64), np.arange(64, 128), indexing="ij")
    pr = np.stack([aa.ravel(), bb.ravel()], 1)
    ref = O.pairs(codes, wl.y[phen], pr)
    c = counts.reshape(-1, 9)
    rel = lambda x, r: np.abs(x - r) / np.maximum(np.abs(r), 1.0)
    rf, rf2 = rel(F.ravel(), ref["F"]), rel(F2.ravel(), ref["F2"])
    out = {"pairs_checked": int(len(pr)), "counts_exact": bool((c == ref["counts"]).all()),
           "f_relerr_max": float(rf.max()), "f2_relerr_max": float(rf2.max()),
           "f_relerr_p50": float(np.percentile(rf, 50)), "f_relerr_p99": float(np.percentile(rf, 99)),
           "f2_relerr_p50": float(np.percentile(rf2, 50)), "f2_relerr_p99": float(np.percentile(rf2, 99)),
           "relerr_policy": "|x - x64| / max(|x64|, 1)"}
    # hit decisions of the block (src/EpiOpenCL.c:251, :487-489) at the workload's thresholds and at the
    # block's own 90th percentiles of F and F2 (so that there are hits to compare); a differing pair
    # must be a tie, |F - t| / t <= 1e-5
    nfac = (ref["counts"] > 0).sum(1)
    nn = ref["counts"].sum(1)
    ok = (nfac >= 2) & (nn - nfac >= 1)
    def hits(f, f2, t1, t2):
        if wl.test == "f":
            return ok & (f > t1)
        return ok & ((f > t1) | (f2 > t2)) & (nfac > 6)
    fin = np.isfinite(ref["F"]) & np.isfinite(ref["F2"])
    levels = [("workload", wl.t1, wl.t2),
              ("block p90", float(np.percentile(ref["F"][fin], 90)), float(np.percentile(ref["F2"][fin], 90)))]
    out["hit_sets"] = []
    for name, t1, t2 in levels:
        g, r = hits(F.ravel(), F2.ravel(), t1, t2), hits(ref["F"], ref["F2"], t1, t2)
        d = g != r
        tie = (np.abs(ref["F"] - t1) <= 1e-5 * t1) | (np.abs(ref["F2"] - t2) <= 1e-5 * t2)
        out["hit_sets"].append({"thresholds": name, "t1": t1, "t2": t2, "hits_gpu": int(g.sum()),
                                "hits_oracle": int(r.sum()), "differ": int(d.sum()),
                                "differ_not_tie": int((d & ~tie).sum())})
    return out


def stress_hit_sets(E, s, wl, p=1e-4):
    """The whole workload scanned by both kernel forms at thresholds for p < 1e-4 (F^-1(1-p; 8 and 4, N-9)),
    where the null alone gives about p * pairs hits per test: the two forms are independent implementations
    (INT8 MFMA on exact fixed-point sums; AND+popcount with FP32 sums) whose FP64 decisions must agree apart
    from pairs within 1e-5 of a threshold."""
    t1 = float(np.float32(E.f_isf(p, 8, wl.nid - 9)))
    t2 = float(np.float32(E.f_isf(p, 4, wl.nid - 9)))
    got = {}
    for form in ("mfma", "plain"):
        r = s.scan(test=wl.test, form=form, t1=t1, t2=t2)
        if r["overflow"]:  # grow the hit buffer to the counted total and scan again
            s.set_hit_capacity(r["nhits"] + 1024)
            r = s.scan(test=wl.test, form=form, t1=t1, t2=t2)
        assert not r["overflow"]
        h = r["hits"]
        got[form] = {(int(a), int(b)): (f, f2) for a, b, f, f2 in zip(h["snp1"], h["snp2"], r["F64"], r["F2_64"])}
    diff = set(got["mfma"]) ^ set(got["plain"])
    def tie(pr):
        f, f2 = got["mfma"].get(pr) or got["plain"][pr]
        return abs(f - t1) <= 1e-5 * t1 or abs(f2 - t2) <= 1e-5 * t2
    return {"p": p, "t1": t1, "t2": t2, "hits_mfma": len(got["mfma"]), "hits_plain": len(got["plain"]),
            "differ": len(diff), "differ_not_tie": sum(not tie(pr) for pr in diff)}


def run_c2(E, s, steps=3):
    """c2: N=4000, M=20000 over 2 chromosomes, 1% missing calls, 2% NaN phenotypes, test i: the plain popcount
    kernel's configuration (every tile is GENERAL), both forms, autotuned schedules, oracle check of each."""
    from epigpu_amd import workloads
    wl = workloads.get("c2").generate()
    s.set_genotypes(wl.words, wl.nid)
    s.set_phenotypes(wl.y)
    s.set_chromosomes(wl.chr_starts, wl.chr_sizes)
    out = {"workload": f"c2: N={wl.nid}, M={wl.nsnp} (2 chr), {wl.pairs} pairs, 1% missing calls, 2% NaN, test i",
           "pairs_per_step": wl.pairs}
    for form in ("plain", "mfma"):
        s.autotune(test=wl.test, form=form, t1=wl.t1, t2=wl.t2, reps=1)
        kms = []
        for i in range(steps + 1):
            r = s.scan(test=wl.test, form=form, t1=wl.t1, t2=wl.t2)
            if i:
                kms.append(r["kernel_ms"])
        km = float(np.mean(kms))
        out[form] = {"kernel_ms": km, "pairs_per_s": wl.pairs / (km * 1e-3), "hits": r["nhits"],
                     "schedule": s.schedule(form), "roofline": roofline(form, wl, km), "check": accuracy(s, wl, form)}
        log(f"[bench c2] {form}: {km:.1f} ms kernel, {out[form]['pairs_per_s']:.3e} pairs/s")
    out["mfma_over_plain"] = out["plain"]["kernel_ms"] / out["mfma"]["kernel_ms"]
    out["stress_hit_sets"] = stress_hit_sets(E, s, wl)
    log(f"[bench c2] stress hit sets: {out['stress_hit_sets']}")
    return out


def run_batch(E, s, steps, warmup, tune=True, plain=True, check=True, sched=None):
    """c4: 64 independent scans of uneven N and M in one captured persistent launch per step."""
    from epigpu_amd import workloads
    bw = workloads.get_batch("c4")
    t0 = time.perf_counter()
    bw.generate()
    log(f"[bench c4] generated {len(bw.scans)} scans, {bw.pairs} pairs in {time.perf_counter() - t0:.1f} s")
    b = E.Batch(s)
    t0 = time.perf_counter()
    for w in bw.scans:
        d = b.add_dataset(w.words, w.nid, w.chr_starts, w.chr_sizes)
        b.add_phenotypes(d, w.y)
    prep_s = time.perf_counter() - t0
    t1, t2 = bw.scans[0].t1, bw.scans[0].t2
    out = {"workload": "c4: 64 scans, N~U[300,6000], M~U[1000,30000] over 1-4 chr, N(0,1) 3% NaN, test i, "
                       f"t1={t1:g}, t2={t2:g}", "scans": len(bw.scans), "pairs_per_step": bw.pairs,
           "prep_ms": prep_s * 1e3}
    if sched:
        b.set_schedule("mfma", **sched)
    first = b.scan(test="i", form="mfma", t1=t1, t2=t2)
    first = b.scan(test="i", form="mfma", t1=t1, t2=t2)
    out["default_schedule"] = dict(b.schedule("mfma"), kernel_ms=first["kernel_ms"])
    log(f"[bench c4] default schedule {out['default_schedule']}")
    if tune:
        t0 = time.perf_counter()
        res = b.autotune(test="i", form="mfma", t1=t1, t2=t2, reps=1)
        out["autotune"] = {"best": res["best"], "best_ms": res["best_ms"], "candidates": len(res["log"]),
                           "tune_s": time.perf_counter() - t0, "log": res["log"]}
        log(f"[bench c4] autotune: best {res['best']} {res['best_ms']:.1f} ms over {len(res['log'])} candidates")
    for i in range(warmup):
        b.scan(test="i", form="mfma", t1=t1, t2=t2)
    kms = []
    t0 = time.perf_counter()
    for i in range(steps):
        st = b.scan(test="i", form="mfma", t1=t1, t2=t2)
        kms.append(st["kernel_ms"])
        log(f"[bench c4] step {i}: {st['kernel_ms']:.1f} ms kernel, {st['nhits']} hits")
    elapsed = time.perf_counter() - t0
    out.update(value=bw.pairs * steps / elapsed, unit="pair-tests/s", ms_per_step=elapsed * 1e3 / steps,
               kernel_ms_per_step=float(np.mean(kms)), schedule=b.schedule("mfma"), grid=st["grid"],
               captures=st["captures"], hits=st["nhits"])
    if check:
        # two scans of the batch against single scans of the same data
        same = True
        for i in (0, len(bw.scans) // 2):
            w = bw.scans[i]
            g = b.hits(i)
            s.set_genotypes(w.words, w.nid)
            s.set_phenotypes(w.y)
            s.set_chromosomes(w.chr_starts, w.chr_sizes)
            one = s.scan(test="i", form="mfma", t1=t1, t2=t2)
            same &= bool((g["hits"]["snp1"] == one["hits"]["snp1"]).all() and
                         (g["hits"]["snp2"] == one["hits"]["snp2"]).all() and len(g["hits"]) == len(one["hits"]))
        out["same_hits_as_single_scans"] = same
    if plain:
        r2 = b.scan(test="i", form="plain", t1=t1, t2=t2)
        r2 = b.scan(test="i", form="plain", t1=t1, t2=t2)
        out["plain_kernel_ms"] = r2["kernel_ms"]
        out["mfma_over_plain"] = r2["kernel_ms"] / out["kernel_ms_per_step"]
        out["plain_same_hit_count"] = r2["nhits"] == out["hits"]
        log(f"[bench c4] plain {r2['kernel_ms']:.1f} ms; mfma/plain {out['mfma_over_plain']:.2f}x")
    b.close()
    return out


def run_multi(E, s, steps, warmup, nsnp=0, check=True, sched=None):
    """c5: one dataset (N=20000, M=200000, 1% missing calls) with 4 phenotypes (N(0,1), Gamma(2,1), t(5),
    0/1 at prevalence 0.2) as 4 scans of one captured persistent launch sharing the genotype planes.
    nsnp > 0 scans only the first nsnp SNPs (a bounded sample of the same workload)."""
    from epigpu_amd import workloads
    wl = workloads.get("c5")
    if nsnp:
        wl.chr_sizes = [nsnp]
    t0 = time.perf_counter()
    wl.generate()
    log(f"[bench c5] generated N={wl.nid} M={wl.nsnp} x {len(wl.y)} phenotypes in {time.perf_counter() - t0:.1f} s")
    b = E.Batch(s)
    t0 = time.perf_counter()
    d = b.add_dataset(wl.words, wl.nid, wl.chr_starts, wl.chr_sizes)
    scans = b.add_phenotypes(d, wl.y)
    prep_s = time.perf_counter() - t0
    if sched:
        b.set_schedule("mfma", **sched)
    tests = wl.pairs * len(scans)
    for i in range(max(1, warmup)):
        st = b.scan(test=wl.test, form="mfma", t1=wl.t1, t2=wl.t2)
        log(f"[bench c5] warmup {i}: {st['kernel_ms']:.1f} ms kernel, {st['nhits']} hits, {st['tiles']} items")
    kms = []
    t0 = time.perf_counter()
    for i in range(steps):
        st = b.scan(test=wl.test, form="mfma", t1=wl.t1, t2=wl.t2)
        kms.append(st["kernel_ms"])
        log(f"[bench c5] step {i}: {st['kernel_ms']:.1f} ms kernel, {st['nhits']} hits")
    elapsed = time.perf_counter() - t0
    kernel_ms = float(np.mean(kms))
    out = {"workload": f"c5: N={wl.nid}, M={wl.nsnp} (1 chr), {wl.pairs} pairs x {len(scans)} phenotypes "
                       f"(N(0,1), Gamma(2,1), t(5), Bernoulli(0.2)), genotype missing {wl.gmiss}, test {wl.test}",
           "pair_tests_per_step": tests, "value": tests * steps / elapsed, "unit": "pair-tests/s",
           "ms_per_step": elapsed * 1e3 / steps, "kernel_ms_per_step": kernel_ms, "prep_ms": prep_s * 1e3,
           "items": st["tiles"], "grid": st["grid"], "schedule": b.schedule("mfma"),
           "hits_per_phenotype": [b.hit_count(sc)[0] for sc in scans]}
    logical = 36.0 * wl.nid * tests
    out["roofline"] = {"bound": "mfma", "achieved": logical / (kernel_ms * 1e-3) / 1e12, "peak": INT8_PEAK_TOPS,
                       "unit": "TFLOP/s"}
    out["roofline"]["frac"] = out["roofline"]["achieved"] / INT8_PEAK_TOPS
    if check:
        s.set_genotypes(wl.words, wl.nid)
        s.set_phenotypes(wl.y)
        s.set_chromosomes(wl.chr_starts, wl.chr_sizes)
        out["check"] = [accuracy(s, wl, "mfma", r) for r in range(len(scans))]
        log(f"[bench c5] oracle check {out['check']}")
    b.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c3")
    ap.add_argument("--form", default="mfma", choices=["mfma", "plain"])
    ap.add_argument("--no-plain", action="store_true", help="skip the plain-form comparison scan")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-check", action="store_true", help="skip the oracle check")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-batch", action="store_true", help="skip the batched c4 measurement")
    ap.add_argument("--no-tune", action="store_true", help="keep the default schedule (no autotuning)")
    ap.add_argument("--batch-steps", type=int, default=3)
    ap.add_argument("--fast-rows", type=int, default=0, help="FAST tile height of the single scan (0: default)")
    ap.add_argument("--schedule", default="", metavar="ROWS,SA,SB[,QUEUES,ORDER]",
                    help="fix the schedule (FAST tile height, super-tile; batch: queues, order) instead of autotuning")
    ap.add_argument("--c5-snps", type=int, default=0, help="c5: scan only the first K SNPs (0: all 200000)")
    ap.add_argument("--gmiss", type=float, default=-1.0,
                    help="single-scan configs: genotype missing-call rate instead of the workload's own (changes the "
                         "workload, which the JSON's config then names)")
    ap.add_argument("--dense-missing", type=int, default=-1, choices=[-1, 0, 1, 2],
                    help="MFMA tiles with missing calls: 0 SPARSE at low rates, 1 GENERAL, 2 SPARSE wherever allowed "
                         "(-1: the schedule's or the autotuner's choice)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run an elementwise copy of MIB MiB (FP32 y = x + 1) for HBM counter calibration")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = args.copy_calibration * (1 << 20) // 4
        x = torch.zeros(n, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            torch.add(x, 1.0, out=y)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_read": 4 * n, "bytes_written": 4 * n, "launches": 3}))
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: the scan runs on one MI355X per process and this build has no "
              "multi-GPU path (pairs would shard as independent tile ranges over several processes). Use --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    import epigpu_amd as E
    from epigpu_amd import workloads

    if args.config in ("c4", "c5"):
        s = E.Scanner(0)
        log(f"[bench] device {s.device_name()}")
        sched = None
        if args.schedule:
            v = [int(x) for x in args.schedule.split(",")]
            sched = dict(zip(("fast_rows", "super_a", "super_b", "queues", "order"), v))
        if args.config == "c4":
            r = run_batch(E, s, args.steps, args.warmup, tune=not (args.no_tune or sched), plain=not args.no_plain,
                          check=not args.no_check, sched=sched)
        else:
            r = run_multi(E, s, args.steps, args.warmup, nsnp=args.c5_snps, check=not args.no_check, sched=sched)
        s.close()
        out = {"metric": METRIC, "value": r["value"], "unit": "pair-tests/s", "n_gpus": 1, "steps": args.steps,
               "warmup": args.warmup, "ms_per_step": r["ms_per_step"], "higher_is_better": True, "scaling": "weak",
               "vs_baseline": None, "dtype": "int8", "data": "synthetic (seeded generator, see epigpu_amd/workloads.py)",
               "config": {"workload": r.pop("workload"), "form": "mfma"}, "detail": r}
        print(json.dumps(out), flush=True)
        return

    wl = workloads.get(args.config)
    if args.gmiss >= 0:
        wl.gmiss = args.gmiss
        wl.name = f"{wl.name} with {args.gmiss:g} missing calls"
    t0 = time.perf_counter()
    wl.generate()
    log(f"[bench] generated {wl.name}: N={wl.nid} M={wl.nsnp} pairs={wl.pairs} in {time.perf_counter() - t0:.1f} s")

    s = E.Scanner(0)
    log(f"[bench] device {s.device_name()}")
    if args.fast_rows:
        s.set_schedule(args.form, fast_rows=args.fast_rows)
    if args.schedule:
        rows, sa, sb = (int(x) for x in args.schedule.split(",")[:3])
        s.set_schedule(args.form, fast_rows=rows, super_a=sa, super_b=sb)
    if args.dense_missing >= 0:
        s.set_schedule(args.form, dense_missing=args.dense_missing)
    t0 = time.perf_counter()
    s.set_genotypes(wl.words, wl.nid)
    s.set_phenotypes(wl.y)
    s.set_chromosomes(wl.chr_starts, wl.chr_sizes)
    prep_s = time.perf_counter() - t0
    log(f"[bench] upload + transcode {prep_s * 1e3:.1f} ms")

    def step(form):
        return s.scan(test=wl.test, form=form, t1=wl.t1, t2=wl.t2)

    tuned = None
    if not args.no_tune and not args.fast_rows and not args.schedule and args.dense_missing < 0:
        t0 = time.perf_counter()
        tuned = s.autotune(test=wl.test, form=args.form, t1=wl.t1, t2=wl.t2, reps=1)
        tuned["tune_s"] = time.perf_counter() - t0
        log(f"[bench] autotune: {tuned['best']} {tuned['best_ms']:.1f} ms over {len(tuned['log'])} candidates "
            f"in {tuned['tune_s']:.1f} s")

    for i in range(args.warmup):
        r = step(args.form)
        log(f"[bench] warmup {i}: {r['kernel_ms']:.1f} ms kernel, {r['nhits']} hits")
    kms = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        r = step(args.form)
        kms.append(r["kernel_ms"])
        log(f"[bench] step {i}: {r['kernel_ms']:.1f} ms kernel, {r['nhits']} hits")
    elapsed = time.perf_counter() - t0
    value = wl.pairs * args.steps / elapsed
    hits = r["hits"]
    found = 0
    if wl.planted_pairs is not None:
        hs = set(zip(hits["snp1"].tolist(), hits["snp2"].tolist()))
        found = sum((int(a), int(b)) in hs for a, b, _, _ in wl.planted_pairs)

    kernel_ms = float(np.mean(kms))
    tiles = s.last_tiles()
    traffic = profiled_traffic(wl.name, args.form, s.schedule(args.form))
    out = {
        "metric": METRIC, "value": value, "unit": "pair-tests/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": elapsed * 1e3 / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "int8" if args.form == "mfma" else "fp32",
        "data": "synthetic (seeded generator, see epigpu_amd/workloads.py)",
        "config": {"workload": f"{wl.name}: N={wl.nid}, M={wl.nsnp} ({len(wl.chr_sizes)} chr), "
                               f"{wl.pairs} pairs, test {wl.test}, t1={wl.t1:g}, t2={wl.t2:g}, "
                               f"MAF U{wl.maf}, genotype missing {wl.gmiss}",
                   "form": args.form, "pairs_per_step": wl.pairs},
        "kernel_ms_per_step": kernel_ms,
        "roofline": roofline(args.form, wl, kernel_ms, traffic,
                             PROFILED.get((wl.name, args.form, s.schedule(args.form)["fast_rows"]), (None,))[0]),
        "hits": int(r["nhits"]), "planted_found": found,
        "prep_ms": prep_s * 1e3, "tiles": tiles,
    }
    if not args.no_plain and args.form == "mfma":
        r2 = step("plain")
        out["plain_kernel_ms"] = r2["kernel_ms"]
        out["plain_pairs_per_s"] = wl.pairs / (r2["kernel_ms"] * 1e-3)
        out["plain_roofline"] = roofline("plain", wl, r2["kernel_ms"])
        out["mfma_over_plain"] = out["plain_kernel_ms"] / kernel_ms
        same = set(zip(r2["hits"]["snp1"].tolist(), r2["hits"]["snp2"].tolist())) == \
            set(zip(hits["snp1"].tolist(), hits["snp2"].tolist()))
        out["plain_same_hits"] = bool(same)
        log(f"[bench] plain form: {r2['kernel_ms']:.1f} ms kernel; mfma/plain = {out['mfma_over_plain']:.2f}x")
    if not args.no_check:
        out["check"] = accuracy(s, wl, args.form)
        log(f"[bench] oracle check: {out['check']}")
        if not args.no_plain and args.form == "mfma":
            out["stress_hit_sets"] = stress_hit_sets(E, s, wl)
            log(f"[bench] stress hit sets: {out['stress_hit_sets']}")
    out["config"]["schedule"] = s.schedule(args.form)
    if tuned:
        out["autotune"] = tuned
    if not args.no_batch:
        out["batch_c4"] = run_batch(E, s, args.batch_steps, 1, plain=not args.no_plain, check=not args.no_check)
        if not args.no_plain and not args.no_check:
            out["c2"] = run_c2(E, s)
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", min(16, os.cpu_count() or 1)))
        out["cpu_baseline"] = cpu_baseline(wl, args.cpu_seconds, threads)
        log(f"[bench] cpu baseline: {out['cpu_baseline']}")
    s.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
