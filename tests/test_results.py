"""CPU: result files of a solve (SimulationResult::save,
core/wavey-solver/src/result.rs:444-461), written by the library and read
back with a restatement of SimulationResult::load (result.rs:463-494).

The format is pinned by the reference's declarations: serde writes struct
fields in declaration order (SimulationResult result.rs:16-38, MicResult
:40-56, Microphone receivers.rs:5-10, RunStats result.rs:80-90,
FdtdSettings fdtd.rs:85-114, with `accuracy` skipped when High and the
`#[serde(skip)]` transfer and late fields absent), and results.bin holds
every microphone's [speaker][freq] transfers as complex f32, microphones in
order, then late energies only when has_late. The reference ships no result
files, so no byte-level fixture exists.
"""
import json

import numpy as np
import pytest

from wavey_amd import fdtd

NB = 8


def load(directory):
    """result.rs:463-494 restated: the JSON metadata, then results.bin read
    in the order save wrote it."""
    meta = json.loads(open(directory / "results.json", "rb").read())
    raw = np.fromfile(directory / "results.bin", dtype=np.float32)
    ns, nf = len(meta["speaker_ids"]), len(meta["freqs"])
    pos = 0
    for m in meta["mics"]:
        n = ns * nf
        m["transfer"] = raw[pos:pos + 2 * n].view(np.complex64).reshape(ns, nf)
        pos += 2 * n
    assert not meta["planes"]
    if meta["has_late"]:
        for m in meta["mics"]:
            m["late"] = raw[pos:pos + ns * NB].reshape(ns, NB)
            pos += ns * NB
    assert pos == raw.size, "results.bin holds more than the counts need"
    return meta


def order(pairs):
    return [k for k, _ in pairs]


def sample(rng, n_mics=3, n_spk=2, nf=5):
    freqs = np.sort(rng.uniform(20.0, 300.0, nf))
    mics = [{"id": 1000 + k, "name": f"mic {k}", "position": rng.uniform(0.0, 5.0, 3)} for k in range(n_mics)]
    mics[-1]["name"] = 'front "row" \\ Café\n'
    tr = rng.normal(size=(n_mics, n_spk, nf)) + 1j * rng.normal(size=(n_mics, n_spk, nf))
    return freqs, [2 ** 40 + 7 + k for k in range(n_spk)], mics, tr


def test_round_trip(tmp_path):
    rng = np.random.default_rng(4)
    freqs, ids, mics, tr = sample(rng)
    stats = {"backend": "MI355X (HIP)", "seconds": 1.25, "images": 0, "paths": 0, "diffracted_paths": 0,
             "receivers": 3, "notes": ["wave solver: 302 x 227 x 102 cells", "fits"]}
    settings = {"crossover_hz": 250.0, "duration_s": 0.0, "points_per_wavelength": 15.0, "source_order": 1,
                "cabinets": True, "accuracy": "high"}
    d = tmp_path / "results" / "run-001"
    fdtd.save_results(d, freqs, ids, mics, tr, stats, fingerprint=2 ** 63 + 5, rt_sabine=np.arange(8) * 0.1,
                      fdtd_settings=settings)
    r = load(d)
    assert r["freqs"] == list(freqs)
    assert r["speaker_ids"] == ids
    for k, m in enumerate(r["mics"]):
        assert m["mic"] == {"id": mics[k]["id"], "name": mics[k]["name"], "position": list(mics[k]["position"])}
        assert m["late_gain"] == [] and m["echograms"] == []
        assert np.array_equal(m["transfer"], tr[k].astype(np.complex64))
    assert r["stats"] == stats
    assert r["fingerprint"] == 2 ** 63 + 5
    assert r["rt_sabine"] == list(np.arange(8) * 0.1) and r["rt_eyring"] == [0.0] * NB
    assert r["has_late"] is False and r["planes"] == []
    assert r["fdtd"] == {k: v for k, v in settings.items() if k != "accuracy"}
    got = fdtd.read_results_transfer(d, len(mics), len(ids), len(freqs))
    assert np.array_equal(got, tr.astype(np.complex64))


def test_field_order_and_types(tmp_path):
    rng = np.random.default_rng(5)
    freqs, ids, mics, tr = sample(rng, n_mics=1, n_spk=1, nf=2)
    fdtd.save_results(tmp_path, freqs, ids, mics, tr, {"backend": "x"},
                      fdtd_settings={"crossover_hz": 200, "accuracy": "draft"})
    text = open(tmp_path / "results.json", "rb").read()
    top = json.loads(text, object_pairs_hook=lambda p: p)
    assert order(top) == ["freqs", "speaker_ids", "mics", "planes", "stats", "fingerprint", "rt_sabine",
                          "rt_eyring", "has_late", "fdtd"]
    d = dict(top)
    mic = d["mics"][0]
    assert order(mic) == ["mic", "late_gain", "echograms"]
    assert order(dict(mic)["mic"]) == ["id", "name", "position"]
    assert order(d["stats"]) == ["backend", "seconds", "images", "paths", "diffracted_paths", "receivers", "notes"]
    assert order(d["fdtd"]) == ["crossover_hz", "duration_s", "points_per_wavelength", "source_order", "cabinets",
                                "accuracy"]
    assert dict(d["fdtd"])["accuracy"] == "draft"
    # Floats stay floats (serde writes 200.0, not 200); ids stay integers.
    r = json.loads(text)
    assert isinstance(r["fdtd"]["crossover_hz"], float) and isinstance(r["stats"]["seconds"], float)
    assert isinstance(r["speaker_ids"][0], int) and isinstance(r["mics"][0]["mic"]["id"], int)
    assert b", " not in text and b": " not in text  # compact, as serde_json::to_vec
    assert (tmp_path / "results.bin").stat().st_size == 1 * 1 * 2 * 8


def test_no_wave_stage_and_non_finite(tmp_path):
    freqs, ids, mics, tr = sample(np.random.default_rng(6), n_mics=2, n_spk=1, nf=3)
    tr[0, 0, 1] = complex(np.nan, np.inf)
    fdtd.save_results(tmp_path, freqs, ids, mics, tr, {"seconds": float("nan")})
    r = load(tmp_path)
    assert r["fdtd"] is None and r["stats"]["seconds"] is None  # serde_json writes NaN as null
    t = r["mics"][0]["transfer"]
    assert np.isnan(t[0, 1].real) and np.isinf(t[0, 1].imag)
    with pytest.raises(fdtd.FdtdError):
        fdtd.read_results_transfer(tmp_path, 3, 1, 3)  # more than the file holds
