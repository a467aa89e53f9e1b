"""The bench JSON line's contract, built from synthetic form records (no GPU)."""

import json
import types

import bench


def _form(name, ps, kernel="k"):
    return {"form": name, "particle_steps_per_s": ps, "ms_per_step": 1.0, "kernel": kernel, "kernel_ms": 0.8,
            "algo_bytes": 2.0e9, "achieved_GBs": 2500.0, "active_particles": 16e6}


def _args(**kw):
    a = dict(steps=50, warmup=5, particles=16_000_000)
    a.update(kw)
    return types.SimpleNamespace(**a)


def test_headline_is_fastest_ab3_form_and_keys_present():
    forms = {"plain_fp32": _form("plain_fp32", 1e9), "fused_fp32": _form("fused_fp32", 1.7e10),
             "fused_bf16": _form("fused_bf16", 1.9e10), "fused_rk2_fp32": _form("fused_rk2_fp32", 5e10)}
    cpu = {"value": 3e7, "unit": "particle-steps/s", "cores": 16, "kind": "port", "sample": "s"}
    # with an accuracy run, the headline is the fastest AB3 form within the bar
    acc = {"steps": 1000, "plain_fp32": {"max": 0.0}, "fused_fp32": {"max": 7e-6}, "fused_bf16": {"max": 0.98}}
    out = bench.build_output(_args(), forms, (1024, 768, 50), cpu, acc, None, None)
    json.dumps(out)
    assert out["value"] == 1.7e10 and out["config"]["form"] == "fused_fp32"  # BF16 misses the bar, RK2 is not AB3
    assert out["forms"]["fused_bf16"]["meets_accuracy_bar"] is False
    acc["fused_bf16"]["max"] = 5e-4
    assert bench.build_output(_args(), forms, (1024, 768, 50), cpu, acc, None, None)["config"]["form"] == "fused_bf16"
    # without one, only FP32-storage forms qualify
    out = bench.build_output(_args(), forms, (1024, 768, 50), cpu, None, None, None)
    assert out["config"]["form"] == "fused_fp32"
    for k in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better",
              "scaling", "vs_baseline", "dtype", "data", "config", "roofline", "cpu_baseline"):
        assert k in out, k
    r = out["roofline"]
    assert r["bound"] == "hbm" and r["peak"] == 8000.0 and r["unit"] == "GB/s"
    assert abs(r["frac"] - 2500.0 / 8000.0) < 1e-12
    assert r["traffic"] is not None  # read from the committed profile of this form and workload
    assert out["dtype"] == "f32" and out["n_gpus"] == 1 and out["vs_baseline"] is None


def test_rk2_only_run_and_other_workloads():
    forms = {"fused_rk2_fp32": _form("fused_rk2_fp32", 1e10)}
    out = bench.build_output(_args(particles=1000), forms, (64, 48, 8), None, None, None, None)
    assert out["config"]["form"] == "fused_rk2_fp32" and out["config"]["workload"].startswith("RK2")
    assert out["roofline"]["traffic"] is None  # no committed profile for this workload
    out = bench.build_output(_args(), {"plain_fp32": _form("plain_fp32", 1e9)}, (1024, 768, 50), None, None,
                             None, None)
    assert out["dtype"] == "f64"


def test_layout_names_match_engine():
    from offline_particles_amd.engine import LAYOUT_NAMES

    assert bench.LAYOUTS == {name: value for value, name in LAYOUT_NAMES.items()}


def test_traffic_scaled_from_nearest_profiled_launch_and_valu_roofline():
    f = _form("fused_fp32", 4e10)
    f.update(steps_per_launch=28, kernel_ms=8.0)
    out = bench.build_output(_args(), {"fused_fp32": f}, (1024, 768, 50), None, None, None, None)
    r = out["roofline"]
    t32 = bench.profiled_traffic("fused_fp32", 16_000_000, (1024, 768, 50), 32)
    assert r["traffic"] == t32 * 28 / 32 and "scaled to 28" in r["traffic_source"]  # nearest profiled: 32
    v = out["roofline_valu"]
    if v is not None:  # needs the committed counter summary of the form
        assert v["bound"] == "valu" and 0 < v["frac"] and v["peak"] == bench.VALU_PEAK_INSTS


def test_batched_only_run_heads_with_the_batching_form():
    batching = {"batched_bf16": {"form": "batched_bf16", "particle_steps_per_s": 3.5e10, "releases": 256}}
    out = bench.build_output(_args(), {}, (1024, 768, 50), None, None, batching, None)
    assert out["value"] == 3.5e10 and out["config"]["form"] == "batched_bf16"
    json.dumps(out)


def test_extra_rows_pass_through_and_cpu_legs_run():
    """The isopycnal and output-event records appear in the line; their CPU legs
    (oracle timings) run on a small workload; adding density leaves the shared
    workload untouched."""
    from offline_particles_amd.workloads import eddy_workload

    iso = {"form": "plain_iso_fp32", "particle_steps_per_s": 1e9}
    outev = {"form": "output_fp32", "particle_outputs_per_s": 1e10}
    out = bench.build_output(_args(), {"plain_fp32": _form("plain_fp32", 1e9)}, (1024, 768, 50), None, None,
                             None, None, iso, outev)
    assert out["isopycnal"] is iso and out["output_events"] is outev
    json.dumps(out)

    wl = eddy_workload(steps=8, nparticles=2000)
    wl_iso = bench.with_density(wl)
    assert "rho" in wl_iso.fields and "rho" not in wl.fields
    c = bench.cpu_baseline(wl_iso, 0.5, 2000, scheme="ab3_iso", max_steps=2)
    assert c["value"] > 0 and "isopycnal" in c["sample"] and c["unit"] == "particle-steps/s"
    o = bench.cpu_output_baseline(wl, 0.5, 2000)
    assert o["value"] > 0 and o["unit"] == "particle-outputs/s"
