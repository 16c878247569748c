"""Copy the reference's committed input bundles and Warp outputs into
tests/golden/ as data fixtures (inputs and expected outputs only).

Run once in a container where the reference tree is mounted:
    python tests/golden/make_fixtures.py /root/reference
Bundles are copied byte for byte so their manifest sha256 still matches.
Large FDTDX arrays are reduced to the entries the tests compare.
"""

import json
import shutil
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent


def copy(src, dst):
    dst.parent.mkdir(parents=True, exist_ok=True)
    shutil.copyfile(src, dst)


def main(root):
    res = Path(root) / "results"
    # Input bundles (manifest + inputs.npz), loadable by Solver(case).
    for case in ("forward50_gpu", "hann_200fs", "gradient20_highest_exact", "forward10_gpu",
                 "hann_10nm", "bend_flow20"):
        for name in ("manifest.json", "inputs.npz"):
            copy(res / "warp_reference" / case / name, HERE / "bundles" / case / name)

    out = HERE / "expected"
    # forward50_gpu: full 3-D Warp fields after 2,098 steps; FDTDX flux traces.
    copy(res / "warp/forward50_gpu/fields.npz", out / "forward50_gpu/warp_fields.npz")
    with np.load(res / "forward50_gpu/arrays.npz") as a:
        np.savez_compressed(out / "forward50_gpu/fdtdx_flux.npz",
                            input=a["input"], output=a["output"])
    # hann20: Warp modal phasors, mid-z field slices and measurements.
    for name in ("mode_in", "mode_out", "mode_reflected"):
        copy(res / f"warp/hann20/{name}_phasors.npz", out / f"hann20/{name}_phasors.npz")
    copy(res / "warp/hann20/fields.npz", out / "hann20/warp_fields_midz.npz")
    r = json.loads((res / "warp/hann20/result.json").read_text())
    (out / "hann20/measurements.json").write_text(
        json.dumps({"warp": r["measurements"], "relative_l2": r["relative_l2"]}, indent=1))
    # gradient20: Warp voxel gradient (input of the pullback) and gradient.
    copy(res / "warp/gradient20/gradient.npz", out / "gradient20/warp_gradient.npz")
    r = json.loads((res / "warp/gradient20/result.json").read_text())
    (out / "gradient20/result.json").write_text(json.dumps(
        {k: r[k] for k in ("measurements", "finite_difference", "relative_l2", "cosine",
                           "history_bytes", "median_seconds")}, indent=1))
    # 24M-cell long-horizon runs (400x400x150 at 10 nm): Warp mid-z field slices
    # after 10,490 (forward10_gpu) and 15,735 (hann10) steps, and hann10's
    # modal measurements.
    for case in ("forward10_gpu", "hann10"):
        copy(res / f"warp/{case}/fields.npz", out / f"{case}/warp_fields_midz.npz")
        r = json.loads((res / f"warp/{case}/result.json").read_text())
        (out / f"{case}/measurements.json").write_text(
            json.dumps({"warp": r["measurements"], "relative_l2": r["relative_l2"],
                        "steps": r["steps"], "shape": r["shape"]}, indent=1))
    # bend_flow20: the flow-plane PhasorDetector (6, 200, 200, 1) at z = 37,
    # mid-z slices and measurements of the 7,868-step run.
    copy(res / "warp/bend_flow20/flow_phasors.npz", out / "bend_flow20/flow_phasors.npz")
    copy(res / "warp/bend_flow20/fields.npz", out / "bend_flow20/warp_fields_midz.npz")
    r = json.loads((res / "warp/bend_flow20/result.json").read_text())
    (out / "bend_flow20/measurements.json").write_text(
        json.dumps({"warp": r["measurements"], "relative_l2": r["relative_l2"]}, indent=1))
    # DrivenUniform observer samples (u050_small: 0.05 spacing, 462 steps).
    copy(res / "octree_guide_controls/u050_small/fields.npz", out / "u050_small/fields.npz")
    r = json.loads((res / "octree_guide_controls/u050_small/result.json").read_text())
    (out / "u050_small/result.json").write_text(json.dumps(
        {k: r[k] for k in ("kind", "spacing", "cells", "problem", "steps", "dt", "dtype",
                           "uniform_layout", "fields_sha256")}, indent=1))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
