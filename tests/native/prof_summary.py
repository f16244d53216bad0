"""Condense rocprofv3 CSV output into the short text summaries kept under
profiles/.

  prof_summary.py stats <dir>/<prefix> [marker] kernel-trace run of bench.py
  prof_summary.py pmc <fetch_csv> <write_csv>   FETCH_SIZE / WRITE_SIZE runs
  prof_summary.py pmcwin <fetch_csv> <write_csv>  the same over one Beyn window
  prof_summary.py mfma <csv>                    MFMA busy, LDS conflicts, waits per GEMM
  prof_summary.py l1 <sq_csv> <ta_csv> <pairs>  address-unit / L1 counters of k_assemble

`stats` prints the whole-run per-kernel table and, from the trace, the last
Beyn window alone (the dispatches from the last table build on), which is the
window bench.py times.  With a marker kernel name, the window is the one that
holds the marker's last dispatch: from the table build before it up to the
next table build (bench.py --symmetry runs its unreduced reference window
after the reduced ones).  `pmc` scales each counter by what it reads for the
plain copy kernel k_copy16 (known bytes, 1 GiB read + 1 GiB written) and
prints HBM bytes per dispatch beside the algorithmic bytes of the GEMM
shapes the probe launches.
"""
import collections
import csv
import sys


def _short(name):
    return name.split("(")[0].replace("void ", "")


def stats(prefix, marker=None):
    rows = list(csv.DictReader(open(prefix + "_kernel_stats.csv")))
    print("whole run (rocprofv3 --stats):")
    print(f"{'total ms':>10} {'calls':>7} {'avg ms':>9}  kernel")
    for r in rows[:24]:
        print(f"{int(r['TotalDurationNs']) / 1e6:10.3f} {int(r['Calls']):7d} "
              f"{float(r['AverageNs']) / 1e6:9.4f}  {_short(r['Name'])}")
    tr = list(csv.DictReader(open(prefix + "_kernel_trace.csv")))
    builds = [i for i, r in enumerate(tr) if "k_build_tables<0>" in r["Kernel_Name"]]
    end = len(tr)
    if marker:
        last = max(i for i, r in enumerate(tr) if marker in r["Kernel_Name"])
        start = max(i for i in builds if i < last)
        end = min([i for i in builds if i > last] + [len(tr)])
    else:
        start = builds[-1]
    sel = tr[start:end]
    agg = collections.defaultdict(lambda: [0, 0.0])
    for r in sel:
        d = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e6
        a = agg[_short(r["Kernel_Name"])]
        a[0] += 1
        a[1] += d
    span = (int(sel[-1]["End_Timestamp"]) - int(sel[0]["Start_Timestamp"])) / 1e6
    busy = sum(v[1] for v in agg.values())
    print(f"\nlast window (trace, dispatches {start}..{end - 1}): span {span:.1f} ms, "
          f"kernel busy {busy:.1f} ms (two streams overlap)")
    print(f"{'total ms':>10} {'calls':>7} {'avg ms':>9}  kernel")
    for k, v in sorted(agg.items(), key=lambda kv: -kv[1][1]):
        print(f"{v[1]:10.3f} {v[0]:7d} {v[1] / v[0]:9.4f}  {k}")
    c64 = [v for k, v in agg.items() if k.startswith("k_gemm_sub<cf")]
    n = sum(v[0] for v in c64)
    ms = sum(v[1] for v in c64)
    print(f"\ncomplex64 GEMM (all tile variants) in the last window: {n} launches, {ms:.2f} ms, "
          f"average {ms / max(n, 1):.4f} ms per launch")


def pmc(fetch_csv, write_csv):
    val = {}
    for path, cname in ((fetch_csv, "FETCH_SIZE"), (write_csv, "WRITE_SIZE")):
        rows = [r for r in csv.DictReader(open(path)) if r["Counter_Name"] == cname]
        val[cname] = [(_short(r["Kernel_Name"]), float(r["Counter_Value"]),
                       (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e6) for r in rows]
    copy_bytes = float(1 << 30)
    f_copy = [v for k, v, _ in val["FETCH_SIZE"] if k == "k_copy16"][0] * 1024
    w_copy = [v for k, v, _ in val["WRITE_SIZE"] if k == "k_copy16"][0] * 1024
    sf, sw = copy_bytes / f_copy, copy_bytes / w_copy
    print(f"copy kernel k_copy16, 1 GiB read + 1 GiB written: FETCH_SIZE {f_copy / 2**30:.4f} GiB, "
          f"WRITE_SIZE {w_copy / 2**30:.4f} GiB")
    print(f"scale factors: FETCH x{sf:.3f}, WRITE x{sw:.3f}\n")
    shapes = {"k_gemm_sub<cf, 128, 128, 16, 2, 2>": (7936, 7936, 256, 4, 8),
              "k_gemm_sub<cd, 128, 64, 16, 2, 2>": None}
    print(f"{'kernel':44} {'ms':>8} {'fetch GB':>9} {'write GB':>9} {'HBM GB':>8} {'alg GB':>8} "
          f"{'GB/s':>7}")
    seen = collections.Counter()
    for (k, fv, ms), (k2, wv, _) in zip(val["FETCH_SIZE"], val["WRITE_SIZE"]):
        if not k.startswith("k_gemm_sub"):
            continue
        seen[k] += 1
        fb, wb = fv * 1024 * sf, wv * 1024 * sw
        alg = None
        if k in shapes and shapes[k]:
            M, N, K, b, es = shapes[k]
            alg = b * es * (M * K + K * N + 2 * M * N)
        elif k.startswith("k_gemm_sub<cd"):
            # probe order: two trailing-update launches, then two residual launches
            M, N, K, b, es = ((7936, 7936, 256, 4, 16) if seen[k] <= 2 else (8192, 768, 8192, 4, 16))
            alg = b * es * (M * K + K * N + 2 * M * N)
        print(f"{k:44} {ms:8.3f} {fb / 1e9:9.3f} {wb / 1e9:9.3f} {(fb + wb) / 1e9:8.3f} "
              f"{(alg or 0) / 1e9:8.3f} {(fb + wb) / 1e6 / ms:7.0f}")


def pmcwin(fetch_csv, write_csv):
    """Window run of pmc_probe.py: a 1 GiB k_copy16 (scale), a 1 MiB k_copy16
    (marker), then one mixed-precision window.  Per kernel family: launches,
    summed HBM bytes (scaled counters) and bytes per launch."""
    per = {}
    for path, cname in ((fetch_csv, "FETCH_SIZE"), (write_csv, "WRITE_SIZE")):
        rows = [r for r in csv.DictReader(open(path)) if r["Counter_Name"] == cname]
        copies = [i for i, r in enumerate(rows) if _short(r["Kernel_Name"]) == "k_copy16"]
        big, mark = copies[0], copies[1]
        scale = float(1 << 30) / (float(rows[big]["Counter_Value"]) * 1024)
        agg = collections.defaultdict(lambda: [0, 0.0])
        for r in rows[mark + 1:]:
            k = _short(r["Kernel_Name"])
            k = "k_gemm_sub<cf> (all tiles)" if k.startswith("k_gemm_sub<cf") else \
                "k_gemm_sub<cd> (all tiles)" if k.startswith("k_gemm_sub<cd") else k
            agg[k][0] += 1
            agg[k][1] += float(r["Counter_Value"]) * 1024 * scale
        per[cname] = (scale, agg)
    sf, af = per["FETCH_SIZE"]
    sw, aw = per["WRITE_SIZE"]
    print(f"scale factors from the 1 GiB copy: FETCH x{sf:.3f}, WRITE x{sw:.3f}")
    print(f"{'launches':>8} {'fetch GB':>9} {'write GB':>9} {'MB/launch':>10}  kernel")
    for k in sorted(af, key=lambda k: -(af[k][1] + aw.get(k, [0, 0])[1])):
        n = af[k][0]
        f, w = af[k][1], aw.get(k, [0, 0.0])[1]
        print(f"{n:8d} {f / 1e9:9.3f} {w / 1e9:9.3f} {(f + w) / 1e6 / max(n, 1):10.3f}  {k}")


def mfma(path, simds=1024, xcds=8):
    """One counter run over pmc_probe.py: SQ_VALU_MFMA_BUSY_CYCLES,
    SQ_LDS_BANK_CONFLICT, SQ_LDS_IDX_ACTIVE, SQ_WAIT_INST_ANY, SQ_WAVE_CYCLES,
    GRBM_GUI_ACTIVE.  GRBM_GUI_ACTIVE is summed over the 8 XCDs, so the
    kernel's active cycles are that sum / 8; MFMA busy = busy cycles /
    (active cycles x 1024 SIMDs)."""
    agg = collections.OrderedDict()
    for r in csv.DictReader(open(path)):
        k = (int(r["Dispatch_Id"]), _short(r["Kernel_Name"]))
        agg.setdefault(k, {})[r["Counter_Name"]] = float(r["Counter_Value"])
    print(f"{'dispatch':>8} {'MFMA busy':>9} {'LDS confl':>9} {'wait/wave':>9}  kernel")
    for (d, k), v in agg.items():
        if not k.startswith("k_gemm_sub"):
            continue
        act = v["GRBM_GUI_ACTIVE"] / xcds
        busy = v["SQ_VALU_MFMA_BUSY_CYCLES"] / (act * simds)
        lds = v["SQ_LDS_BANK_CONFLICT"] / max(v["SQ_LDS_IDX_ACTIVE"], 1.0)
        wait = v["SQ_WAIT_INST_ANY"] / max(v["SQ_WAVE_CYCLES"], 1.0)
        print(f"{d:8d} {100 * busy:8.1f}% {100 * lds:8.1f}% {100 * wait:8.1f}%  {k}")


def l1(sq_csv, ta_csv, pairs, cus=256, ses=32):
    """Counters of the k_assemble dispatch, summed over the device's
    instances: TA and TCP per CU (cus), SQ per shader engine (ses, the
    wave/busy cycle counters in units of 4 cycles), GRBM_GUI_ACTIVE over the
    8 XCDs.  `pairs` = entry pairs of the dispatch (N(N+1)/2 per node)."""
    agg = collections.defaultdict(float)
    grbm = []
    for path in (sq_csv, ta_csv):  # one run each; both also read GRBM_GUI_ACTIVE
        for r in csv.DictReader(open(path)):
            if "k_assemble<" in r["Kernel_Name"]:
                if r["Counter_Name"] == "GRBM_GUI_ACTIVE":
                    grbm.append(float(r["Counter_Value"]))
                else:
                    agg[r["Counter_Name"]] += float(r["Counter_Value"])
    cyc = sum(grbm) / len(grbm) / 8  # per XCD: the kernel's cycles
    print(f"kernel cycles (GRBM_GUI_ACTIVE / 8 XCDs): {cyc:.4g}")
    print(f"TA busy per CU:                 {agg['TA_TA_BUSY'] / cus / cyc:.1%} of the cycles")
    print(f"TA stalled by the L1 per CU:    {agg['TA_ADDR_STALLED_BY_TC_CYCLES'] / cus / cyc:.1%}")
    print(f"L1 pending-miss stalls per CU:  {agg['TCP_PENDING_STALL_CYCLES'] / cus / cyc:.1%}")
    print(f"L1 accesses per entry pair:     {agg['TCP_TOTAL_ACCESSES'] / pairs:.1f}")
    print(f"L1 accesses per vector load:    {agg['TCP_TOTAL_ACCESSES'] / agg['SQ_INSTS_VMEM_RD']:.1f}")
    print(f"L1 accesses per CU per cycle:   {agg['TCP_TOTAL_ACCESSES'] / cus / cyc:.2f}")
    print(f"CU cycles per entry pair:       {cus * cyc / pairs:.1f}")
    print(f"L1 -> L2 read requests:         {agg['TCP_TCC_READ_REQ']:.4g}")
    print(f"VALU busy per SIMD:             "
          f"{agg['SQ_ACTIVE_INST_VALU'] * 4 / (4 * cus) / cyc:.1%}")
    print(f"waves resident per CU:          {agg['SQ_WAVE_CYCLES'] * 4 / cus / cyc:.1f}")


if __name__ == "__main__":
    if sys.argv[1] == "l1":
        l1(sys.argv[2], sys.argv[3], float(sys.argv[4]))
    elif sys.argv[1] == "mfma":
        mfma(sys.argv[2])
    elif sys.argv[1] == "stats":
        stats(sys.argv[2], sys.argv[3] if len(sys.argv) > 3 else None)
    elif sys.argv[1] == "pmcwin":
        pmcwin(sys.argv[2], sys.argv[3])
    else:
        pmc(sys.argv[2], sys.argv[3])
