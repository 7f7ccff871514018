"""CPU: the scan driver's host logic with a stand-in sector solver (no device).

* a scan whose k-th sector raises leaves a checkpoint with the finished
  sectors; a rerun on the same path loads it and solves nothing, and a rerun
  with ``overwrite=True`` solves only the sectors not yet in the results
  (SRE:2462-2522);
* the flux scan writes one checkpoint per twist under the reference's file
  name (SRE:2571-2600), refreshes a copy of the catalog (never the caller's)
  to each twist's group, and skips twists whose checkpoint exists
  (SRE:2672-2677);
* the checkpoint file name against a hand-written expected string.
"""

from __future__ import annotations

import os
from fractions import Fraction

import numpy as np
import pytest

from cases import make_case
from realspace_exatdiagonalization_py_amd import symmetry_resolved_ed as sre
from realspace_exatdiagonalization_py_amd.groups import build_irrep_list


class StandInCatalog:
    """Picklable stand-in for the device catalog: records refreshes and copies."""

    def __init__(self, origin=None):
        self.origin = origin
        self.groups = []

    def clone(self):
        return StandInCatalog(origin=self)


@pytest.fixture
def stand_in(monkeypatch):
    solved = []
    fail_at = {"n": None}

    def solve(label, ed_data, nev=5, **kw):
        if fail_at["n"] is not None and len(solved) == fail_at["n"]:
            raise RuntimeError("sector failed")
        idx = sre._find_irrep_idx(ed_data, label)
        phase = np.asarray(ed_data.symmetry_group.operations[1].perm_phases)
        ed_data.ed_scan_res[idx] = (np.array([float(idx), float(np.angle(phase).sum())]), np.zeros((1, 1)))
        ed_data.sector_dims[idx] = idx + 1
        solved.append((label, id(ed_data)))

    def rephase(catalog, group):
        catalog.groups.append(group)
        return catalog

    monkeypatch.setattr(sre, "ed_scan_at_irrep_matrixfree", solve)
    monkeypatch.setattr(sre, "_rephase_catalog", rephase)
    return solved, fail_at


def _data():
    model, ne, G = make_case("hal23")
    irreps = build_irrep_list(G, model.lattice)
    return sre.Symmetry_Resolved_ED_Data(model, ne, Fraction(ne, model.lattice.n_site), G, irreps,
                                         StandInCatalog(), [0] * len(irreps))


def test_failed_sector_rerun_skips_finished(tmp_path, stand_in):
    solved, fail_at = stand_in
    data = _data()
    path = str(tmp_path / "scan.pkl")
    fail_at["n"] = 3
    with pytest.raises(RuntimeError, match="sector failed"):
        sre.ed_scan(data, nev=2, mode="matrixfree", checkpoint_path=path)
    assert sorted(data.ed_scan_res) == [0, 1, 2]
    assert sorted(sre.load_checkpoint(path).ed_scan_res) == [0, 1, 2]
    fail_at["n"] = None
    # same path, no overwrite: the checkpoint is loaded and nothing is solved
    fresh = _data()
    n0 = len(solved)
    assert sre.ed_scan(fresh, nev=2, mode="matrixfree", checkpoint_path=path) == [path]
    assert len(solved) == n0 and sorted(fresh.ed_scan_res) == [0, 1, 2]
    # overwrite: finished sectors (in the results) are skipped, the rest solved
    sre.ed_scan(data, nev=2, mode="matrixfree", checkpoint_path=path, overwrite=True)
    again = [lab for lab, _ in solved[n0:]]
    finished = {data.irrep_list[i].label for i in (0, 1, 2)}
    assert not finished & set(again)
    assert len(again) == len(data.irrep_list) - 3
    assert sorted(sre.load_checkpoint(path).ed_scan_res) == list(range(len(data.irrep_list)))


def test_flux_scan_checkpoints_and_skip(tmp_path, stand_in):
    solved, _ = stand_in
    data = _data()
    original = data.orbit_catalog
    model = data.second_quantized_model
    d = str(tmp_path / "flux")
    thetas = [0.0, 0.25, 0.5]
    labels = [ir.label for ir in data.irrep_list[:2]]
    paths = sre.ed_scan(data, nev=2, mode="matrixfree", scanned_sectors=labels, checkpoint_dir=d,
                        twisted_phases_over_2π_list=thetas)
    assert len(paths) == 3 and original.groups == []  # the caller's catalog is never refreshed
    for θ, p in zip(thetas, paths):
        assert os.path.dirname(p) == d
        assert os.path.basename(p) == sre.ed_scan_checkpoint_filename(model, [θ, 0.0], data.filling_fraction)
        loaded = sre.load_checkpoint(p)
        assert sorted(loaded.ed_scan_res) == [0, 1]
        # results were computed with this twist's group: the x-translation's phases
        phase = np.asarray(loaded.symmetry_group.operations[1].perm_phases)
        assert loaded.ed_scan_res[0][0][1] == pytest.approx(float(np.angle(phase).sum()))
        assert loaded.orbit_catalog.origin is not None  # the scan's copy, at this twist
        np.testing.assert_array_equal(np.asarray(loaded.orbit_catalog.groups[-1].operations[1].perm_phases),
                                      phase)
    n0 = len(solved)
    again = sre.ed_scan(data, nev=2, mode="matrixfree", scanned_sectors=labels, checkpoint_dir=d,
                        twisted_phases_over_2π_list=thetas + [0.75])
    assert again[:3] == paths and len(again) == 4
    assert len(solved) - n0 == len(labels)  # only the new twist was solved
    with pytest.raises(ValueError, match="Unknown ED scan mode"):
        sre.ed_scan(data, mode="dense")


def test_checkpoint_filename_format():
    model, ne, G = make_case("hal23")
    got = sre.ed_scan_checkpoint_filename(model, [0.25, 0.0], Fraction(1, 4))
    params = {k: round(v, 3) if isinstance(v, (int, float)) else v for k, v in model.params.items()}
    want = (f"{model.tb_model.model_name}_{model.lattice.sample_size}_ν_graph=1_4_"
            f"twisted_phases_over_2π=[0.25, 0.0]_params={params}.pkl")
    assert got == want
    assert got.startswith(model.tb_model.model_name + "_[2, 3]_ν_graph=1_4_twisted_phases_over_2π=[0.25, 0.0]")
