"""The convergence record bench.py writes from --converge outputs and the
time to 0.5 cm^-1 it derives from it: the point count a form needs is the
smallest from which every checkpoint of every seed has all levels within
0.5 cm^-1 of the reference (mean of the gradient estimator's seeds), and a
merge keeps the record's earlier checkpoints."""
import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402

REF = np.array([100.0, 200.0, 300.0])
PR = {"config": "toy", "cv_t": 3, "cv_nn": 4}


def lines(rows):
    return "\n".join(json.dumps(r) for r in rows) + "\n"


def row(form, seed, points, dev, points_t=None):
    return {"form": form, "seed": seed, "points": points, "points_t": points_t, "accumulate_s": points / 100.0,
            "train_s": 0.0, "levels_cm1": list(REF + np.array(dev))}


@pytest.fixture
def record(tmp_path, monkeypatch):
    monkeypatch.setattr(bench, "CONV_DIR", str(tmp_path))
    stein = [row("stein_mixed", s, 64, [d, -d, 0.0]) for s, d in ((0, 0.1), (1, -0.1))]
    # seed 1 is within 0.5 from 20 points on; seed 0 enters the band at 16,
    # leaves it at 24 and stays in from 32: at 24 a later row of the same
    # point count (seed 1) is within the band, the earlier one is not
    mixed = [row("mixed", 0, 8, [-0.9, 0, 0]), row("mixed", 0, 16, [-0.4, 0, 0]), row("mixed", 0, 24, [-0.6, 0, 0]),
             row("mixed", 0, 32, [-0.2, 0, 0]), row("mixed", 1, 16, [0, -0.7, 0]), row("mixed", 1, 20, [0, 0.45, 0]),
             row("mixed", 1, 24, [0, -0.45, 0]), row("mixed", 1, 32, [0, 0.1, -0.3])]
    cv = [row("cv", 0, 16, [0.2, 0, 0], points_t=8), row("cv", 0, 16, [3.0, 0, 0], points_t=16)]
    f = tmp_path / "runs.jsonl"
    f.write_text(lines(stein + mixed + cv + [mixed[-1]]))  # a repeated checkpoint is dropped
    bench.write_convergence(dict(PR), [str(f)], 10)
    return tmp_path


def test_points_required_over_all_seeds(record):
    rec = json.load(open(record / "convergence_toy_10states.json"))
    np.testing.assert_allclose(rec["reference"]["levels_cm1"], REF)
    mixed = rec["forms"]["mixed"]
    assert len(mixed["checkpoints"]) == 8
    assert mixed["points_required"] == 32
    assert sorted(mixed["dev_at_required_cm1"]) == ["0", "1"]
    np.testing.assert_allclose(mixed["dev_at_required_cm1"]["1"], [0.0, 0.1, -0.3], atol=1e-9)
    # only the control variate rows with the preset's 2^cv_t true points
    cv = rec["forms"]["cv"]
    assert cv["points_t"] == 8 and cv["points_required"] == 16 and len(cv["checkpoints"]) == 1


def test_merge_keeps_checkpoints(record, tmp_path):
    g = tmp_path / "more.jsonl"
    g.write_text(lines([row("mixed", 2, 32, [0, 0, 0.6]), row("mixed", 2, 48, [0, 0, 0.2])]))
    bench.write_convergence(dict(PR), [str(g)], 10, merge=True)
    rec = json.load(open(record / "convergence_toy_10states.json"))
    mixed = rec["forms"]["mixed"]
    assert len(mixed["checkpoints"]) == 10
    assert mixed["points_required"] == 48
    np.testing.assert_allclose(rec["reference"]["levels_cm1"], REF)
    # merging the same lines again changes nothing
    before = (record / "convergence_toy_10states.json").read_text()
    bench.write_convergence(dict(PR), [str(g)], 10, merge=True)
    assert (record / "convergence_toy_10states.json").read_text() == before


def test_time_to_half_wavenumber(record):
    res = {"fp64": {"points": 100, "t_acc": 10.0, "t_lev": 0.5},
           "mixed": {"points": 100, "t_acc": 1.0, "t_lev": 0.5},
           "cv": {"t_acc": 2.0, "t_lev": 0.5, "train_s": 3.0}}
    out = bench.time_to_half_wavenumber(dict(PR), res, 10)
    assert out["fp64"]["s"] == pytest.approx(32 / 10.0 + 0.5)
    assert out["mixed"]["s"] == pytest.approx(32 / 100.0 + 0.5)
    # one-product surrogate (the default): the correction's 2^cv_t true
    # points are each a mixed-form point and a surrogate point
    rate_nn = (16 + 8) / (2.0 - 8 / 100.0)
    assert out["cv"]["s"] == pytest.approx(3.0 + (16 + 8) / rate_nn + 8 / 100.0 + 0.5)
    assert out["speedup_cv_vs_fp64"] == pytest.approx(out["fp64"]["s"] / out["cv"]["s"])
    # three-product surrogate: the correction is one mixed-form pass
    out3 = bench.time_to_half_wavenumber(dict(PR, cv_terms=3), res, 10)
    rate3 = 16 / (2.0 - 8 / 100.0)
    assert out3["cv"]["s"] == pytest.approx(3.0 + 16 / rate3 + 8 / 100.0 + 0.5)


def test_summary_without_reference_fails_cleanly(tmp_path, monkeypatch):
    monkeypatch.setattr(bench, "CONV_DIR", str(tmp_path))
    f = tmp_path / "runs.jsonl"
    f.write_text(lines([row("mixed", 0, 8, [0.1, 0, 0])]))
    with pytest.raises(SystemExit):
        bench.write_convergence(dict(PR), [str(f)], 10, merge=True)
