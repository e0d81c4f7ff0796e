"""The counter decode behind profiles/pmc_traffic_*.json
(ntustcfdlab_amd/rocprof.py), on synthetic counter summaries: request
sizes, the calibration check and the per-launch entries the bench reads."""
import csv

from ntustcfdlab_amd import rocprof


def counters(rd128, rd64, rd32, wr64, wr32):
    reads = {"TCC_EA0_RDREQ_sum": rd128 + rd64 + rd32, "TCC_EA0_RDREQ_64B_sum": rd64,
             "TCC_EA0_RDREQ_32B_sum": rd32}
    writes = {"TCC_EA0_WRREQ_sum": wr64 + wr32, "TCC_EA0_WRREQ_64B_sum": wr64}
    return reads, writes


def summaries(copy_err=0.0):
    mib = 1 << 20
    # a 64 MiB copy: 128-byte reads, 64-byte writes (plus an error factor)
    cr, cw = counters(64 * mib // 128 * (1 + copy_err), 0, 0, 64 * mib // 64, 0)
    name0 = "void ntust::k_rb_sweep32<0, 2>(ntust::KGrid, ...)"
    name1 = "void ntust::k_rb_sweep32<1, 2>(ntust::KGrid, ...)"
    r0, w0 = counters(1000, 10, 4, 200, 6)
    r1, w1 = counters(600, 0, 2, 100, 0)
    reads = {"copies": [dict(cr, dispatch="1")],
             "counters": {name0: dict(r0, dispatches=5), name1: dict(r1, dispatches=5)}}
    writes = {"copies": [dict(cw, dispatch="2")],
              "counters": {name0: dict(w0, dispatches=5), name1: dict(w1, dispatches=5)}}
    return reads, writes


def test_decode_and_validation():
    reads, writes = summaries()
    t = rocprof.traffic(8, reads, writes, 400.0)
    assert t["validated"] and t["sweep_cap"] == 400.0 and t["cells"] == 512
    cal = t["calibration_copy"][0]
    assert cal["bytes_each_way"] == 64 << 20 and cal["read_err"] == 0.0 and cal["write_err"] == 0.0
    k0 = t["kernels"]["k_rb_sweep32<0, 2>"]
    assert k0["read_bytes"] == 128 * 1000 + 64 * 10 + 32 * 4
    assert k0["write_bytes"] == 64 * 200 + 32 * 6
    assert k0["dispatches"] == 5 and k0["read_B_per_cell"] == k0["read_bytes"] / 512
    k1 = t["kernels"]["k_rb_sweep32<1, 2>"]
    head = t["per_launch"]["k_rb_sweep32<C>"]
    assert head["read_bytes"] == (k0["read_bytes"] + k1["read_bytes"]) / 2
    assert head["write_bytes"] == (k0["write_bytes"] + k1["write_bytes"]) / 2


def test_a_decode_off_by_more_than_5_percent_is_not_validated():
    assert rocprof.traffic(8, *summaries(0.04))["validated"]
    assert not rocprof.traffic(8, *summaries(0.06))["validated"]


def test_summarize_reads_trace_and_counter_csv(tmp_path):
    with open(tmp_path / "x_kernel_trace.csv", "w", newline="") as fh:
        w = csv.DictWriter(fh, ["Kernel_Name", "Start_Timestamp", "End_Timestamp"])
        w.writeheader()
        for t0, t1 in ((1000, 3000), (5000, 9000)):
            w.writerow({"Kernel_Name": "k", "Start_Timestamp": t0, "End_Timestamp": t1})
    with open(tmp_path / "x_counter_collection.csv", "w", newline="") as fh:
        w = csv.DictWriter(fh, ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"])
        w.writeheader()
        # two instances of one counter in one dispatch add up
        for did, name, val in (("1", "k_stream_copy", 3.0), ("1", "k_stream_copy", 4.0),
                               ("2", "k", 10.0), ("3", "k", 20.0)):
            w.writerow({"Dispatch_Id": did, "Kernel_Name": name, "Counter_Name": "C",
                        "Counter_Value": val})
    s = rocprof.summarize(str(tmp_path))
    assert s["trace"]["k"]["calls"] == 2 and s["trace"]["k"]["avg_us"] == 3.0
    assert s["span_ms"] == (9000 - 1000) / 1e6
    assert s["counters"]["k"]["dispatches"] == 2 and s["counters"]["k"]["C"] == 15.0
    assert s["copies"] == [{"dispatch": "1", "C": 7.0}]
