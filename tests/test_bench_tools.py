"""Host-side helpers of bench.py: the launch-choice keys of the traffic
record and the mapping of per-dispatch counter rows to launch choices."""
import csv
import importlib.util
import json
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _csv(path, counter, rows):
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Dispatch_Id", "Kernel_Name", "Counter_Name",
                                          "Counter_Value"])
        w.writeheader()
        for d, kern, v in rows:
            w.writerow(dict(Dispatch_Id=d, Kernel_Name=kern, Counter_Name=counter,
                            Counter_Value=v))


def test_tuning_keys():
    b = _bench()
    assert b.tuning_key(None) == "default"
    assert b.tuning_key(dict(order=3, bucket_bits=0, xcd=2, tile=1)) == "order3_bb0_xcd2_t16"
    assert b.tuning_key(dict(order=1, bucket_bits=0, xcd=0, tile=0, chunk_log2=18)) == \
        "order1_bb0_xcd0_c18"
    tun = b.sweep_tunings("tuner")
    keys = [b.tuning_key(t) for t in tun]
    # order 0 and 1 without bucket edges, 2 and 3 with three each; two tiles; six XCD modes
    assert len(keys) == len(set(keys)) == (1 + 1 + 3 + 3) * 2 * 6
    assert b.product_kernel("batched", "bf16x4", dict(tile=1), 50) == "k_product_t16"
    assert b.product_kernel("batched", "bf16x4", dict(tile=1), 120) == "k_product_mx"
    assert b.product_kernel("batched", "fp32", dict(tile=1), 50) == "k_product_mx"


def test_sweep_dispatches_map_to_launch_choices(tmp_path, monkeypatch):
    b = _bench()
    monkeypatch.setattr(b, "TRAFFIC_FILE", str(tmp_path / "traffic.json"))
    sweep = "3,0,2,1;0,0,1,0"
    t16 = "void lk::k_product_t16<4, 4>(lk::ObsDev, int)"
    mx = "void lk::k_product_mx<64, 4>(lk::ObsDev, int)"
    other = "void lk::k_search<3>(lk::LocDev)"
    (tmp_path / "f").mkdir()
    (tmp_path / "w").mkdir()
    # rows out of dispatch order, with another kernel between
    _csv(tmp_path / "f" / "c.csv", "FETCH_SIZE", [(9, mx, 300.0), (4, t16, 100.0), (5, other, 7.0)])
    _csv(tmp_path / "w" / "c.csv", "WRITE_SIZE", [(4, t16, 10.0), (9, mx, 30.0), (5, other, 1.0)])
    b.summarize_pmc(str(tmp_path / "f" / "c.csv"), str(tmp_path / "w" / "c.csv"), "c3",
                    "batched", "bf16x4", 1000, None, sweep)
    tr = json.load(open(tmp_path / "traffic.json"))
    a = tr["c3/batched/bf16x4/order3_bb0_xcd2_t16"]
    c = tr["c3/batched/bf16x4/order0_bb0_xcd1"]
    assert a["kernel"] == "k_product_t16" and c["kernel"] == "k_product_mx"
    assert a["bytes_per_dispatch"] == round(1024 * (100.0 * b.FETCH_SCALE + 10.0 * b.WRITE_SCALE))
    assert c["bytes_per_dispatch"] == round(1024 * (300.0 * b.FETCH_SCALE + 30.0 * b.WRITE_SCALE))
    # a dispatch count that does not match the list is refused
    _csv(tmp_path / "w" / "c.csv", "WRITE_SIZE", [(4, t16, 10.0)])
    with pytest.raises(SystemExit):
        b.summarize_pmc(str(tmp_path / "f" / "c.csv"), str(tmp_path / "w" / "c.csv"), "c3",
                        "batched", "bf16x4", 1000, None, sweep)
    # so is a dispatch of the other tile's kernel where this one should be
    _csv(tmp_path / "w" / "c.csv", "WRITE_SIZE", [(4, mx, 10.0), (9, t16, 30.0)])
    with pytest.raises(SystemExit):
        b.summarize_pmc(str(tmp_path / "f" / "c.csv"), str(tmp_path / "w" / "c.csv"), "c3",
                        "batched", "bf16x4", 1000, None, sweep)
