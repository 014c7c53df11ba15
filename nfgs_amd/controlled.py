"""Controlled evaluation of a reconstructed field against a reference.

The evaluator's interface (NPZ keys, configuration keys, result keys and
error messages) is the reference's (nfgs/controlled.py:1-122), so existing
input files, configurations and score tables carry over.  Everything else
is organised here as data plus stages:

* ``VOLUME`` and ``BOUNDARY`` are declared NPZ layouts; ``Layout.read``
  applies the declared component and coordinate rules in order and raises
  the first one that fails.
* ``ScoringBox`` parses and checks the z-y-x scoring box.
* ``Evaluation`` runs three metric stages over one checked input trio:
  fidelity to the reference (FD2 current and divergence of the full volume,
  box sums), the outward bottom-field comparison against the observed input
  and against the clean reference, and the bottom-plane residuals.

All arithmetic runs on the GPU in FP64: the FD2 curl and divergence and the
box sums (nfgs_amd.scores), the Gaussian separation and the outward-field
errors (nfgs_amd.cci).  Reading files, contract checks and assembling the
result are host work.
"""
from __future__ import annotations

import argparse
import hashlib
import json
from dataclasses import dataclass
from pathlib import Path
from typing import Callable

import numpy as np
import torch

from . import cci
from . import scores

COMPONENTS = ("Bx", "By", "Bz")
UNIT_KEYS = ("field_unit", "length_unit")


# ---------------------------------------------------------------------------
# NPZ layouts and their rules
# ---------------------------------------------------------------------------

def _increasing(c: np.ndarray) -> bool:
    return c.ndim == 1 and len(c) >= 3 and bool(np.all(np.isfinite(c))) and \
        bool(np.all(np.diff(c) > 0))


def _uniform(c: np.ndarray) -> bool:
    steps = np.diff(c)
    return bool(np.allclose(steps, steps[0], rtol=1e-7, atol=1e-12))


# (predicate on one coordinate array, message), applied in this order
COORDINATE_RULES: tuple[tuple[Callable[[np.ndarray], bool], str], ...] = (
    (_increasing, "{path}: {axis} must have at least 3 increasing coordinates"),
    (_uniform, "{path}: nonuniform {axis} is not supported by FD2/GS contract"),
)


@dataclass(frozen=True)
class Layout:
    """One NPZ layout: three field components whose array shape is the
    coordinate lengths in reverse axis order, plus the unit strings."""
    axes: tuple

    def required(self) -> set:
        return set(COMPONENTS) | set(self.axes) | set(UNIT_KEYS)

    def read(self, path) -> dict:
        with np.load(path, allow_pickle=False) as npz:
            absent = self.required() - set(npz.files)
            if absent:
                raise ValueError(f"{path}: missing coordinate/unit contract {absent}")
            rec = {name: np.asarray(npz[name], dtype=np.float64)
                   for name in COMPONENTS + self.axes}
            rec.update((u, str(npz[u].item())) for u in UNIT_KEYS)
        shape = tuple(len(rec[a]) for a in self.axes[::-1])
        bad = next((k for k in COMPONENTS
                    if rec[k].shape != shape or not np.all(np.isfinite(rec[k]))), None)
        if bad is not None:
            raise ValueError(f"{path}: invalid {bad} shape or nonfinite data")
        for axis in self.axes:
            for ok, message in COORDINATE_RULES:
                if not ok(rec[axis]):
                    raise ValueError(message.format(path=path, axis=axis))
        return rec


VOLUME = Layout(("x", "y", "z"))
BOUNDARY = Layout(("x", "y"))


def load(path, boundary: bool = False) -> dict:
    """Read one NPZ of the contract (components, coordinates, units) and
    check it (nfgs/controlled.py:25-44)."""
    return (BOUNDARY if boundary else VOLUME).read(path)


def match(reference: dict, candidate: dict, boundary: bool = False) -> None:
    """Units equal and coordinates identical: no resampling is ever implied
    (nfgs/controlled.py:47-52)."""
    unit = next((u for u in UNIT_KEYS if reference[u] != candidate[u]), None)
    if unit is not None:
        raise ValueError(f"Unit mismatch: {unit}")
    axes = (BOUNDARY if boundary else VOLUME).axes
    axis = next((a for a in axes if not np.array_equal(reference[a], candidate[a])), None)
    if axis is not None:
        raise ValueError(f"Coordinate mismatch: {axis}; no implicit resampling is allowed")


def digest(path) -> str:
    """SHA-256 of a file's bytes."""
    return hashlib.sha256(Path(path).read_bytes()).hexdigest()


sha = digest  # the reference's name (nfgs/controlled.py:21-22)


# ---------------------------------------------------------------------------
# Scoring box
# ---------------------------------------------------------------------------

@dataclass(frozen=True)
class ScoringBox:
    spec: list          # as written in the configuration
    slices: tuple       # (z, y, x) slices

    @classmethod
    def parse(cls, spec, shape) -> "ScoringBox":
        """Three unstrided, non-empty slices inside `shape`."""
        slices = tuple(slice(*s) for s in spec)
        if len(slices) != 3:
            raise ValueError("Three z,y,x scoring slices are required")
        for n, s in zip(shape, spec):
            step = s[2] if len(s) > 2 else 1
            if not (0 <= s[0] < s[1] <= n) or step not in (None, 1):
                raise ValueError("Scoring box is outside the field or is strided")
        return cls(spec, slices)

    @property
    def window(self):
        """The y/x part of the box."""
        return self.slices[1:]


# ---------------------------------------------------------------------------
# Outward bottom-field errors
# ---------------------------------------------------------------------------

def cci_errors(target, candidate, slices, cutoffs) -> dict:
    """Relative error of the outward bottom field on the y/x window of
    `slices`: in L2 over the window, and as the mean of |c - b| / |b| over
    the pixels whose |b| exceeds each cutoff fraction of max |b|
    (nfgs/controlled.py:55-68).  Computed by nfgs_outward_errors."""
    cuts = list(cutoffs)
    sums = cci.outward_sums(target, candidate, tuple(slices)[1:], cuts, eps=0.0)
    per_cut = {}
    for k, cut in enumerate(cuts):
        per_cut[str(cut)] = {"E_rel_Bgt": sums.mean_rel(k), "retained_pixels": sums.kept[k],
                             "retained_fraction": sums.kept[k] / sums.pixels}
    return {"E_rel_Bgt_L2": sums.l2(), "CCI_pixels_total": sums.pixels,
            "CCI_cutoff_results": per_cut, "E_rel_Bgt": per_cut[str(cuts[0])]["E_rel_Bgt"]}


# ---------------------------------------------------------------------------
# The evaluation
# ---------------------------------------------------------------------------

# fixed entries of every result (nfgs/controlled.py:98-113)
RESULT_CONSTANTS = {
    "current_operator": "common_fd2_float64_full_volume_then_crop",
    "CCI_target": "actual_observed_input_full_footprint_then_crop",
    "diagnostics_not_computed": ["FastQSL", "twist", "six_face_energy", "helicity"],
    "controlled_acquisition_verified": False,
}


class Evaluation:
    """One evaluation: a configuration and the three checked inputs."""

    def __init__(self, reference_path, field_path, boundary_path, config_path):
        self.paths = {"reference": reference_path, "field": field_path,
                      "boundary": boundary_path, "config": config_path}
        self.config = json.loads(Path(config_path).read_text())
        if self.config.get("schema_version") != 1:
            raise ValueError("Unsupported evaluation config")
        self.ref = load(reference_path)
        self.sol = load(field_path)
        self.obs = load(boundary_path, boundary=True)
        match(self.ref, self.sol)
        match(self.ref, self.obs, boundary=True)
        if self.ref["z"][0] != 0:
            raise ValueError("Photospheric z=0 must be the first plane")
        if any(self.sol[k].shape != self.ref[k].shape for k in COMPONENTS):
            raise ValueError("Field shape mismatch")
        self.box = ScoringBox.parse(self.config["scoring_slices_zyx"], self.ref["Bx"].shape)
        self.spacing = tuple(float(np.diff(self.ref[a])[0]) for a in ("x", "y", "z"))
        self.current = None

    # -- stage 1: agreement with the reference, sigma_J, divergence --------
    def fidelity(self) -> dict:
        self.current, div = scores.curl_and_divergence(self.sol, self.spacing)
        # the divergence is normalised by 6 / h_eff with
        # 1 / h_eff = (1/dx + 1/dy + 1/dz) / 3 (isotropic-grid convention)
        h_eff = 3.0 / sum(1.0 / h for h in self.spacing)
        return scores.compute_metrics(self.ref, self.sol, self.current, div, self.box.slices,
                                      h_eff)

    # -- stage 2: outward bottom field vs the observed input and the clean reference
    def outward(self) -> dict:
        cuts = self.config["CCI_cutoff_fractions"]
        splitter = cci.make_bgt_helper_solver(nx=len(self.ref["x"]), ny=len(self.ref["y"]),
                                              dx=self.spacing[0], dy=self.spacing[1],
                                              dz=self.spacing[2], dtype=torch.float64,
                                              **self.config["gaussian_separation"])
        with torch.inference_mode():
            parts = {}
            for name, planes in (("observed", tuple(self.obs[k] for k in COMPONENTS)),
                                 ("candidate", tuple(self.sol[k][0] for k in COMPONENTS)),
                                 ("clean", tuple(self.ref[k][0] for k in COMPONENTS))):
                parts[name] = cci.outward_field(splitter, *planes).clone()
        out = cci_errors(parts["observed"], parts["candidate"], self.box.slices, cuts)
        out["clean_reference_CCI"] = cci_errors(parts["clean"], parts["candidate"],
                                                self.box.slices, cuts)
        return out

    # -- stage 3: bottom-plane residuals against the input -----------------
    def bottom(self) -> dict:
        sy, sx = self.box.window
        dbz = self.sol["Bz"][0, sy, sx] - self.obs["Bz"][sy, sx]
        # the input's own FD2 Jz: the boundary repeated over three planes
        stack = {k: np.stack([self.obs[k]] * 3) for k in COMPONENTS}
        jz_in = scores.curl(stack, self.spacing)["Jz"][0]
        djz = self.current["Jz"][0, sy, sx] - jz_in[sy, sx]
        return {"bottom_Bz_RMSE_to_input": float(np.sqrt(np.mean(dbz * dbz))),
                "bottom_Jz_RMSE_to_input_fd2": float(np.sqrt(np.mean(djz * djz)))}

    STAGES = ("fidelity", "outward", "bottom")

    def run(self) -> dict:
        metrics = {}
        for stage in self.STAGES:
            metrics.update(getattr(self, stage)())
        result = {"schema_version": 1, "case": self.config["case"],
                  "method": self.config["method"],
                  "current_operator": RESULT_CONSTANTS["current_operator"],
                  "CCI_target": RESULT_CONSTANTS["CCI_target"],
                  "spacing_xyz": self.spacing, "scoring_slices_zyx": self.box.spec,
                  "field_unit": self.ref["field_unit"], "length_unit": self.ref["length_unit"],
                  "source_sha256": {k: digest(p) for k, p in self.paths.items()},
                  "metrics": metrics}
        result.update((k, RESULT_CONSTANTS[k]) for k in
                      ("diagnostics_not_computed", "controlled_acquisition_verified"))
        return result


def evaluate(reference_path, field_path, boundary_path, config_path) -> dict:
    """Scalar diagnostics of a field against a reference and the observed
    boundary (nfgs/controlled.py:71-113)."""
    return Evaluation(reference_path, field_path, boundary_path, config_path).run()


def main(argv=None) -> None:
    """Command line: --reference --field --boundary --config --output
    (nfgs/controlled.py:116-122)."""
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
