"""The per-kernel roofline report (unap_amd/kernel_report.py) on a small
synthetic kernel trace: the load-shape instantiations of one kernel body are
counted as one kernel, the report starts at the last solve's kStateInit, and
the per-level table assigns Chebyshev launches coarse to fine."""
import json

from unap_amd import kernel_report


def _trace(tmp_path, names_ns):
    rows = ["Kernel_Name,Start_Timestamp,End_Timestamp"]
    t = 0
    for name, d in names_ns:
        rows.append(f'"{name}",{t},{t + d}')
        t += d + 10
    p = tmp_path / "trace.csv"
    p.write_text("\n".join(rows) + "\n")
    return p


def test_shapes_merge_and_levels(tmp_path):
    cheb = "void unap_hip::kRows<unap_hip::BCheby<float, {s}> >(unap_hip::BCheby<float, {s}>, int, unap_hip::Skip)"
    init = "void unap_hip::kStateInit(unap_hip::KState*)"
    # an earlier solve that must be ignored, then one V-cycle of a 3-level
    # hierarchy (2 coarse levels: level 1 smoothed, level 2 coarsest):
    # 4 sweeps on level 1 (shape 1), then 4 on the finest (shape 0)
    seq = [(init, 5), (cheb.format(s=0), 999999), (init, 5)]
    seq += [(cheb.format(s=1), 1000)] * 4 + [(cheb.format(s=0), 4000)] * 4
    tr = _trace(tmp_path, seq)
    bench = tmp_path / "bench.json"
    bench.write_text("progress line\n" + json.dumps({"hierarchy": [[10**8, 3 * 10**8], [25 * 10**6, 9 * 10**7],
                                                                   [60, 150]],
                                                     "iterations": 1}) + "\n")
    out = kernel_report.report(str(tr), str(bench))
    lines = out.splitlines()
    row = [ln for ln in lines if ln.startswith("| kRows<BCheby<float> >")]
    assert len(row) == 1, out
    cells = [c.strip() for c in row[0].strip("|").split("|")]
    assert cells[1] == "8"                      # both shapes, one name
    assert abs(float(cells[2]) - 0.02) < 1e-9   # 4*1000 + 4*4000 ns
    # algorithmic bytes: 4 sweeps on the finest and on level 1, FP32
    b = 4 * sum((6 * 4 + 4) * n + (2 * 4 + 8) * f for n, f in [(10**8, 3 * 10**8), (25 * 10**6, 9 * 10**7)])
    assert abs(float(cells[4]) - b / 1e9) < 1e-2
    lvl = {ln.split("|")[1].strip(): ln for ln in lines if ln.startswith("| ") and ln.split("|")[1].strip().isdigit()}
    assert "| 4.0 |" in lvl["0"] and "| 1.0 |" in lvl["1"], out


def test_time_table_last_solve(tmp_path):
    init = "void unap_hip::kStateInit(unap_hip::KState*)"
    lvl = "void unap_hip::kSweepLevel<double, double, 0>(unap_hip::SweepArgs<double, double>, int, int, unap_hip::Skip)"
    seq = [(init, 5), (lvl, 500000), (init, 5), (lvl, 2000), (lvl, 4000)]
    out = kernel_report.time_table(str(_trace(tmp_path, seq)))
    row = [ln for ln in out.splitlines() if "kSweepLevel" in ln]
    assert len(row) == 1, out
    cells = [c.strip() for c in row[0].strip("|").split("|")]
    assert cells[1] == "2" and abs(float(cells[3]) - 3.0) < 1e-9  # two launches, mean 3 us


def test_one_off_cast_counted_per_launch():
    # the FP64 -> FP32 source cast runs once a solve when the Krylov update
    # supplies the FP32 residual: its bytes follow its launches, not the
    # iteration count
    lv = [(1000, 3000), (250, 900), (60, 150)]
    name = "void unap_hip::kRows<unap_hip::BCopy<double, float> >(unap_hip::BCopy<double, float>, int, unap_hip::Skip)"
    assert kernel_report.alg_bytes(name, lv, 83, 1) == 12 * 1000
    assert kernel_report.alg_bytes(name, lv, 83) == 83 * 12 * 1000


def test_restriction_sub_steps_per_level():
    # with pair steps above 2 every level but the last is reached in
    # ceil(pair_steps / 2) restriction launches; the last level takes the
    # launches left over in each V-cycle
    name = "void unap_hip::kRowsXcd<unap_hip::BRestrict<float, float> >(unap_hip::BRestrict<float, float>, int)"
    seg = [{"Kernel_Name": name, "Start_Timestamp": "0", "End_Timestamp": "1"}] * 12
    assert kernel_report.restrict_order(seg, 3, 2, 4) == [0, 0, 1, 1, 2, 2]
    assert kernel_report.restrict_order(seg, 3, 2, 2) == [0, 1, 2, 2, 2, 2]
    assert kernel_report.restrict_order(seg[:6], 3, 2, 2) == [0, 1, 2]
    assert kernel_report.restrict_order(seg[:10], 3, 2, 4) == [0, 0, 1, 1, 2]
    # summed per level and V-cycle
    durs = [{"Kernel_Name": name, "Start_Timestamp": "0", "End_Timestamp": str(d)} for d in (10, 20, 3, 4, 1, 2) * 2]
    got = kernel_report._by_group(durs, "BRestrict<float, float>", [0, 0, 1, 1, 2, 2])
    assert got == {0: 30.0, 1: 7.0, 2: 3.0}
