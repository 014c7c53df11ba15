"""Controlled evaluation of a reconstructed field (scalar diagnostics).

Same contract, names and results as the reference's nfgs/controlled.py:1-113:
NPZ inputs hold Bx/By/Bz in z, y, x order (y, x for the observed boundary),
1-D x/y(/z) coordinates and scalar field_unit / length_unit strings; the
current and divergence are the common FD2 operators on the full volume, then
cropped to the scoring box; the CCI target is the outward Gaussian part of the
actual observed boundary, separated on its full footprint, then cropped.

The FD2 curl and divergence, the box sums and the Gaussian separation run on
the GPU in FP64 (nfgs_amd.scores, nfgs_amd.cci); loading, contract checks and
the plane-sized CCI means are host work, as in the reference.
"""
from __future__ import annotations

import argparse
import hashlib
import json
from pathlib import Path

import numpy as np
import torch

from . import cci
from . import scores

KEYS = ("Bx", "By", "Bz")
UNITS = ("field_unit", "length_unit")


def sha(path) -> str:
    """SHA-256 of a file's bytes (controlled.py:21-22)."""
    return hashlib.sha256(Path(path).read_bytes()).hexdigest()


def load(path, boundary: bool = False) -> dict:
    """Read and check one NPZ of the contract (controlled.py:25-44): the
    three components, the coordinates (x, y and, for a volume, z: at least 3,
    finite, increasing and uniform) and the two unit strings."""
    axes = ("x", "y") if boundary else ("x", "y", "z")
    with np.load(path, allow_pickle=False) as data:
        missing = set(KEYS + axes + UNITS) - set(data.files)
        if missing:
            raise ValueError(f"{path}: missing coordinate/unit contract {missing}")
        out = {k: np.asarray(data[k], dtype=np.float64) for k in KEYS + axes}
        out.update({k: str(data[k].item()) for k in UNITS})
    shape = tuple(len(out[k]) for k in reversed(axes))
    for k in KEYS:
        if out[k].shape != shape or not np.all(np.isfinite(out[k])):
            raise ValueError(f"{path}: invalid {k} shape or nonfinite data")
    for k in axes:
        c = out[k]
        steps = np.diff(c) if c.ndim == 1 else None
        if c.ndim != 1 or len(c) < 3 or not np.all(np.isfinite(c)) or not np.all(steps > 0):
            raise ValueError(f"{path}: {k} must have at least 3 increasing coordinates")
        if not np.allclose(steps, steps[0], rtol=1e-7, atol=1e-12):
            raise ValueError(f"{path}: nonuniform {k} is not supported by FD2/GS contract")
    return out


def match(reference: dict, candidate: dict, boundary: bool = False) -> None:
    """Same units and identical coordinates, no resampling (controlled.py:47-52)."""
    for k in UNITS:
        if reference[k] != candidate[k]:
            raise ValueError(f"Unit mismatch: {k}")
    for k in ("x", "y") if boundary else ("x", "y", "z"):
        if not np.array_equal(reference[k], candidate[k]):
            raise ValueError(f"Coordinate mismatch: {k}; no implicit resampling is allowed")


def cci_errors(target, candidate, slices, cutoffs) -> dict:
    """E_rel of the outward bottom field on the y/x window of `slices`, in L2
    and as pixel means over |b^>| above each cutoff fraction of its maximum
    (controlled.py:55-68)."""
    _, sy, sx = slices
    b = np.asarray(target)[:, sy, sx]
    B = np.asarray(candidate)[:, sy, sx]
    mag = np.linalg.norm(b, axis=0)
    diff = np.linalg.norm(B - b, axis=0)
    total = float(np.linalg.norm(b))
    res = {"E_rel_Bgt_L2": float(np.linalg.norm(B - b) / total) if total > 0 else None,
           "CCI_pixels_total": int(mag.size), "CCI_cutoff_results": {}}
    for cut in cutoffs:
        keep = mag > float(cut) * float(mag.max())
        res["CCI_cutoff_results"][str(cut)] = {
            "E_rel_Bgt": float(np.mean(diff[keep] / mag[keep])) if keep.any() else None,
            "retained_pixels": int(keep.sum()), "retained_fraction": float(keep.mean())}
    res["E_rel_Bgt"] = res["CCI_cutoff_results"][str(cutoffs[0])]["E_rel_Bgt"]
    return res


def _scoring_box(config, shape):
    spec = config["scoring_slices_zyx"]
    box = tuple(slice(*s) for s in spec)
    if len(box) != 3:
        raise ValueError("Three z,y,x scoring slices are required")
    for n, s in zip(shape, spec):
        start, stop = s[:2]
        step = s[2] if len(s) > 2 else 1
        if not 0 <= start < stop <= n or step not in (None, 1):
            raise ValueError("Scoring box is outside the field or is strided")
    return spec, box


def evaluate(reference_path, field_path, boundary_path, config_path) -> dict:
    """Scalar diagnostics of a field against a reference and the observed
    boundary (controlled.py:71-113)."""
    config = json.loads(Path(config_path).read_text())
    if config.get("schema_version") != 1:
        raise ValueError("Unsupported evaluation config")
    ref, field, obs = load(reference_path), load(field_path), load(boundary_path, True)
    match(ref, field)
    match(ref, obs, True)
    if ref["z"][0] != 0:
        raise ValueError("Photospheric z=0 must be the first plane")
    if any(field[k].shape != ref[k].shape for k in KEYS):
        raise ValueError("Field shape mismatch")
    spec, box = _scoring_box(config, ref["Bx"].shape)
    spacing = tuple(float(np.diff(ref[k])[0]) for k in ("x", "y", "z"))

    j, div = scores.curl_and_divergence(field, spacing)
    # 6 / h_eff = 2 (1/dx + 1/dy + 1/dz), the isotropic convention
    h_eff = 3.0 / sum(1.0 / h for h in spacing)
    metrics = scores.compute_metrics(ref, field, j, div, box, h_eff)

    helper = cci.make_bgt_helper_solver(nx=len(ref["x"]), ny=len(ref["y"]), dx=spacing[0],
                                        dy=spacing[1], dz=spacing[2], dtype=torch.float64,
                                        **config["gaussian_separation"])
    cuts = config["CCI_cutoff_fractions"]
    target = cci.bottom_plane_bgt_nfgs(helper, {k: obs[k][None] for k in KEYS})
    candidate = cci.bottom_plane_bgt_nfgs(helper, field)
    metrics.update(cci_errors(target, candidate, box, cuts))
    # the clean-reference CCI, reported apart from the fidelity to the input
    clean = cci.bottom_plane_bgt_nfgs(helper, ref)
    metrics["clean_reference_CCI"] = cci_errors(clean, candidate, box, cuts)
    _, sy, sx = box
    metrics["bottom_Bz_RMSE_to_input"] = float(
        np.sqrt(np.mean((field["Bz"][0, sy, sx] - obs["Bz"][sy, sx]) ** 2)))
    jobs = scores.curl({k: np.repeat(obs[k][None], 3, axis=0) for k in KEYS}, spacing)["Jz"][0]
    metrics["bottom_Jz_RMSE_to_input_fd2"] = float(
        np.sqrt(np.mean((j["Jz"][0, sy, sx] - jobs[sy, sx]) ** 2)))
    paths = {"reference": reference_path, "field": field_path, "boundary": boundary_path,
             "config": config_path}
    return {"schema_version": 1, "case": config["case"], "method": config["method"],
            "current_operator": "common_fd2_float64_full_volume_then_crop",
            "CCI_target": "actual_observed_input_full_footprint_then_crop",
            "spacing_xyz": spacing, "scoring_slices_zyx": spec,
            "field_unit": ref["field_unit"], "length_unit": ref["length_unit"],
            "source_sha256": {k: sha(v) for k, v in paths.items()},
            "metrics": metrics,
            "diagnostics_not_computed": ["FastQSL", "twist", "six_face_energy", "helicity"],
            "controlled_acquisition_verified": False}


def main(argv=None) -> None:
    """Command line of controlled.py:116-122."""
    p = argparse.ArgumentParser(description=__doc__)
    for name in ("reference", "field", "boundary", "config", "output"):
        p.add_argument("--" + name, type=Path, required=True)
    a = p.parse_args(argv)
    result = evaluate(a.reference, a.field, a.boundary, a.config)
    a.output.parent.mkdir(parents=True, exist_ok=True)
    a.output.write_text(json.dumps(result, indent=2, allow_nan=False) + "\n")
    print(json.dumps({"output": str(a.output), "method": result["method"],
                      "metrics": result["metrics"]}, indent=2))


if __name__ == "__main__":
    main()
