"""Writes tests/golden/reference_goldens.json from the reference's own test files.

Run in the build container (the reference checkout is read as text, never
executed):  python tests/golden/extract_reference_goldens.py

Sources (paths relative to the reference checkout):
  crates/indicatrix/tests/raytracer_tests/golden_regression.rs:47-233
      trace_spectral_ray XYZ bits (Diamond/Sapphire x 4 rays), the
      sample_studio_environment bits and evaluate_gem_optical_metrics bits.
  crates/indicatrix/src/optics/raytracer/environment/tests.rs:786-1160
      BASELINE_BITS: 9 presets x 64 Fibonacci directions x 3 wavelengths.
"""
import json
import os
import re
import sys

REF = "/root/reference/crates/indicatrix"
HERE = os.path.dirname(os.path.abspath(__file__))


def strip_comments(text):
    return "\n".join(line.split("//")[0] for line in text.splitlines())


def hexes(text):
    text = strip_comments(text)
    return [int(h.replace("_", ""), 16) for h in re.findall(r"0x([0-9a-fA-F_]+)", text)]


def golden_regression():
    src = open(os.path.join(REF, "tests/raytracer_tests/golden_regression.rs")).read()
    cases_txt = src[src.index("let cases: [GoldenCase; 4]"):src.index("// Every row is traced")]
    bits = hexes(cases_txt)
    assert len(bits) == 24, len(bits)
    # The ray, seed, preset and light pose of each row, as the table states them
    # (golden_regression.rs:47-130); "normalize" marks a direction the test
    # normalizes at run time.
    rows = [
        dict(origin=[0.0, 2.5, 0.0], dir=[0.0, -1.0, 0.0], normalize=False, seed=1337,
             preset="RingLights", light_yaw=0.85, light_pitch=0.95),
        dict(origin=[0.0, 2.5, 0.0], dir=[0.0, -1.0, 0.0], normalize=False, seed=42,
             preset="Incandescent", light_yaw=0.0, light_pitch=1.2),
        dict(origin=[0.3, 2.5, 0.1], dir=[-0.05, -1.0, 0.02], normalize=True, seed=9001,
             preset="DarkSpotlight", light_yaw=2.1, light_pitch=0.4),
        dict(origin=[0.0, 10.0, 0.0], dir=[0.0, -1.0, 0.0], normalize=False, seed=7,
             preset="Daylight", light_yaw="PI", light_pitch=0.3),
    ]
    for i, r in enumerate(rows):
        r["diamond_bits"] = bits[6 * i:6 * i + 3]
        r["sapphire_bits"] = bits[6 * i + 3:6 * i + 6]
    env_txt = src[src.index("let env_expected"):src.index("// The inputs go through")]
    env_bits = hexes(env_txt)
    assert len(env_bits) == 12
    metrics_txt = src[src.index("m.extinction_pct.to_bits(),"):]
    metric_bits = hexes(metrics_txt)[:5]
    return dict(
        source="crates/indicatrix/tests/raytracer_tests/golden_regression.rs",
        cut="standard_round_brilliant", max_bounces=12, exposure=1.0,
        hero_rand="hash_u32(seed) / 4294967295", traces=rows,
        environment=dict(preset="RingLights", exposure=1.0, light_yaw=0.85, light_pitch=0.95,
                         dirs=[[0.0, 1.0, 0.0], [0.5, 0.5, 0.5], [-0.3, 0.2, 0.9], [1.0, 0.0, 0.0]],
                         normalize=[False, True, True, False], lambdas=[450.0, 550.0, 650.0],
                         bits=[env_bits[3 * i:3 * i + 3] for i in range(4)]),
        optical_metrics=dict(material="Diamond", cam_yaw=0.0, cam_pitch=0.45,
                             environment="RingLights.studio(1.0, 0.85, 0.95)",
                             fields=["brilliance_pct", "fire_index", "scintillation_pct",
                                     "windowing_pct", "extinction_pct"], bits=metric_bits))


def environment_baseline():
    src = open(os.path.join(REF, "src/optics/raytracer/environment/tests.rs")).read()
    start = src.index("const BASELINE_BITS: [[u32; 192]; 9] = [")
    body = src[start + len("const BASELINE_BITS: [[u32; 192]; 9] = ["):]
    body = body[:body.index("];\n")]
    rows = []
    for chunk in body.split("],")[:9]:
        vals = [int(v) for v in re.findall(r"\b(\d+)\b", chunk)]
        rows.append(vals)
    assert all(len(r) == 192 for r in rows), [len(r) for r in rows]
    return dict(
        source="crates/indicatrix/src/optics/raytracer/environment/tests.rs",
        order=["Daylight", "Incandescent", "RingLights", "DarkSpotlight", "IsoHemisphere",
               "LightTent", "DaylightDome", "UvLamp365", "UvLamp395"],
        exposure=1.2, light_yaw=0.4, light_pitch=0.35, lambdas=[450.0, 550.0, 650.0],
        max_ulp=2,
        directions="k in 0..64: phi=(k+0.5)*(PI*(3-sqrt(5))); y=fma(k+0.5, -(2/64), 1); "
                   "r=sqrt(max(fma(y,-y,1),0)); dir=(r cos phi, y, r sin phi)",
        bits=rows)


def main():
    if not os.path.isdir(REF):
        sys.exit("reference checkout not present; the committed JSON is the fixture")
    out = dict(golden_regression=golden_regression(), environment_baseline=environment_baseline())
    path = os.path.join(HERE, "reference_goldens.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
