"""The v606 drop-in (gisapr_amd.search) run from files, against the oracle."""
import ast

import numpy as np
import pytest

from oracle import search as osr
from oracle import splat as osp

pytestmark = pytest.mark.gpu


def _parse(path):
    recs = {}
    for line in open(path):
        item = ast.literal_eval(line)
        recs.setdefault(item[1], {})[item[0]] = item
    return recs


@pytest.mark.parametrize("name", ["box32", "box64"])
def test_run_search_from_files(hip_device, tmp_path, name):
    import torch
    from gisapr_amd import mrc, search, star, synthetic
    wl = synthetic.WORKLOADS[name]
    d = synthetic.make_workload(name)
    print("[case] run_search files", name, flush=True)
    imgs = d["images"].numpy()
    mrc.write_mrc(tmp_path / "parts.mrcs", imgs)
    star.write_particle_star(tmp_path / "parts.star", str(tmp_path / "parts.mrcs"), d["rot"], d["tilt"], d["psi"],
                             d["dfu"], d["dfv"], d["dfa"])
    star.write_direction_star(tmp_path / "dirs.star", d["dir_rot"], d["dir_tilt"])
    vol = osp.density_from_coords(d["coords"], d["numbers"], wl.box, wl.apix, wl.res)
    mrc.write_mrc(tmp_path / "model.mrc", vol)
    argv = ["--i", "x", "--p", str(tmp_path / "parts.star"), "--o", str(tmp_path / "out.txt"),
            "--ang", str(tmp_path / "dirs.star"), "--mrc", str(tmp_path / "model.mrc"),
            "--oriboxsize", str(wl.box), "--newboxsize", str(wl.box), "--apix", str(wl.apix),
            "--kk", str(wl.kk), "--psiStep", str(wl.psi_step), "--doLocalSearch",
            "--localRange", str(wl.local_range), "--ccgCropSize", str(wl.crop),
            "--maskRadius", str(wl.mask_radius), "--maskEdge", str(wl.mask_edge)]
    args = search.create_search_parser().parse_args(argv)
    metrics = search.run_search(args)
    assert set(metrics) >= {"search_s", "projection_s", "setup_s", "execution_s"}
    got = _parse(tmp_path / "out.txt")
    # the tensor-input form (worker's search_from_pdb) gives the same records
    args2 = search.create_search_parser().parse_args(argv[:5] + [str(tmp_path / "out2.txt")] + argv[6:])
    search.run_search(args2, model_volume=torch.from_numpy(vol).to(hip_device))
    assert _parse(tmp_path / "out2.txt") == got
    parts = {"images": imgs, "rot": d["rot"], "tilt": d["tilt"], "psi": d["psi"],
             "dfu": d["dfu"], "dfv": d["dfv"], "dfa": d["dfa"]}
    want = osr.run_search(vol.astype(np.float64), d["dir_rot"], d["dir_tilt"], parts, apix=wl.apix,
                          new_box=wl.box, psi_step=wl.psi_step, local_range=wl.local_range, crop=wl.crop,
                          kk=wl.kk, mask_radius=wl.mask_radius, mask_edge=wl.mask_edge, dtype=np.float64)
    table = star.read_particle_table(str(tmp_path / "parts.star"))
    lines = [r.line_index for r in table.records]
    assert sorted(got) == sorted(lines[w["line"]] for w in want)
    for w in want:
        g = got[lines[w["line"]]]
        cc, psi, raw = g["cc: "], g["psi: "], g["rawCC: "]
        assert (cc[5], cc[6], cc[8], cc[9]) == (w["rot"], w["tilt"], w["xshift"], w["yshift"])
        assert psi[2] == w["psi_fake"] and psi[3] == pytest.approx(w["psi_prior"])
        assert raw[2] == pytest.approx(w["rawcc"], rel=1e-4, abs=1e-3)
        assert cc[2] == pytest.approx(w["z"], rel=1e-3, abs=1e-3)
        assert cc[7] == pytest.approx(w["angdist"], abs=1e-9) and cc[3] == w["l3"] and cc[4] == w["accu"]


def test_v606_namespace_mirrors(hip_device):
    """PackedCTFTorch, IntegratedProjector, ParticleFourierPreparer and
    correlate_templates_and_particle keep the v606 tensor contracts."""
    import torch
    from gisapr_amd import host, search, synthetic
    from oracle import fourier as ofr, particles as opt, project as opj
    print("[case] v606 namespace mirrors", flush=True)
    dev = hip_device
    N, n, C = 32, 32, 8
    rng = np.random.default_rng(17)
    # CTF
    ctf = search.PackedCTFTorch(n, 2.0, 2.7, 300.0, 3.0, dev)
    recs = [search.ParticleRecord(0, "", 0, 0, 0, 15000.0, 13000.0, 35.0),
            search.ParticleRecord(1, "", 0, 0, 0, 22000.0, 21000.0, 120.0)]
    got = ctf.batch(recs, True).cpu().numpy()
    octf = opt.PackedCTF(n, 2.0, 2.7, 300.0, 3.0)
    for r, g in zip(recs, got):
        np.testing.assert_allclose(g, octf.one(r.defocus_u, r.defocus_v, r.defocus_angle), atol=2e-4)
    # projector (pre-shifted bank)
    layout = search.PackedFourierLayout(n, C, dev)
    mask = torch.from_numpy(host.generate_soft_mask((N, N), 12, 3)).to(dev)
    fsc = layout.centered_full_to_packed(torch.ones((n, n), device=dev))
    pts = rng.integers(10, 22, (20, 3))
    vol = osp.density_from_coords((pts - N // 2) * 1.0, np.full(20, 6.0), N, 1.0, 2.0)
    angles = search.ModelAngles([10.0, 70.0], [30.0, 100.0], [0.0, 0.0], host.rotation_matrices_t([10.0, 70.0], [30.0, 100.0]))
    proj = search.IntegratedProjector(dev, N, n, layout, mask, fsc)
    bank = proj.build_templates(torch.from_numpy(vol).to(dev), angles).cpu().numpy()
    want = opj.build_templates(vol.astype(np.float64), angles.rot, angles.tilt, n, mask.cpu().numpy(),
                               fsc.cpu().numpy(), np.float64)
    assert np.linalg.norm(bank - want) / np.linalg.norm(want) < 1e-4
    # particle preparation with an explicit packed CTF
    eff = [0.5, 7.5, 15.0]
    bankpsi = search.PsiGridBank(N, dev, eff)
    prep = search.ParticleFourierPreparer(N, n, dev, layout, mask, bankpsi)
    img = rng.standard_normal((N, N)).astype(np.float32)
    c = ctf.batch(recs[:1], True)[0]
    pf = prep.prepare(torch.from_numpy(img).to(dev), eff, c).cpu().numpy()
    pw = opt.prepare_particle(img.astype(np.float64), eff, n, mask.cpu().numpy(), c.cpu().numpy(), np.float64)
    assert np.linalg.norm(pf - pw) / np.linalg.norm(pw) < 1e-4
    # correlation on pre-shifted operands
    sc, ys, xs = search.correlate_templates_and_particle(torch.from_numpy(bank).to(dev), torch.from_numpy(pf).to(dev),
                                                         layout, search.CorrelationWorkspace(dev))
    s2, y2, x2 = ofr.correlate(want, pw, C, np.float64)
    np.testing.assert_allclose(sc.cpu().numpy(), s2, rtol=1e-3, atol=1e-3 * np.abs(s2).max())
    assert np.array_equal(ys.cpu().numpy(), y2) and np.array_equal(xs.cpu().numpy(), x2)
    recs3 = search.make_result_records(search.ParticleRecord(5, "", 10.0, 30.0, 3.0), [0.0, 7.5, 15.0],
                                       angles.rot, angles.tilt, [0, 1], sc, ys, xs, C, 1.0)
    assert [r[0] for r in recs3] == ["cc: ", "psi: ", "rawCC: "] and recs3[2][2] == pytest.approx(float(s2.max()), rel=1e-3)


def test_calibration_bank_round_trip(hip_device, tmp_path):
    """--orientations-out ranks each particle's directions (best over psi,
    Top-K); --orientations-in then scores only those and finds the same best
    direction, psi, shift and rawCC; both projection modes give the same
    records; a restricted bank cannot be recalibrated."""
    import json
    import torch
    from gisapr_amd import mrc, orientation_bank, search, star, synthetic
    wl = synthetic.WORKLOADS["box32"]
    d = synthetic.make_workload("box32")
    print("[case] calibration bank", flush=True)
    mrc.write_mrc(tmp_path / "parts.mrcs", d["images"].numpy())
    star.write_particle_star(tmp_path / "parts.star", str(tmp_path / "parts.mrcs"), d["rot"], d["tilt"], d["psi"],
                             d["dfu"], d["dfv"], d["dfa"])
    star.write_direction_star(tmp_path / "dirs.star", d["dir_rot"], d["dir_tilt"])
    vol = osp.density_from_coords(d["coords"], d["numbers"], wl.box, wl.apix, wl.res)
    mrc.write_mrc(tmp_path / "model.mrc", vol)
    base = ["--i", "x", "--p", str(tmp_path / "parts.star"), "--ang", str(tmp_path / "dirs.star"),
            "--mrc", str(tmp_path / "model.mrc"), "--oriboxsize", str(wl.box), "--newboxsize", str(wl.box),
            "--apix", str(wl.apix), "--kk", str(wl.kk), "--psiStep", str(wl.psi_step), "--doLocalSearch",
            "--localRange", str(wl.local_range), "--ccgCropSize", str(wl.crop), "--maskRadius", str(wl.mask_radius),
            "--maskEdge", str(wl.mask_edge), "--keep-orientations", "2"]
    parse = search.create_search_parser().parse_args
    full = search.run_search(parse(base + ["--o", str(tmp_path / "full.txt"), "--projection_mode", "full"]))
    m = search.run_search(parse(base + ["--o", str(tmp_path / "cal.txt"), "--orientations-out",
                                        str(tmp_path / "bank.json")]))
    assert _parse(tmp_path / "cal.txt") == _parse(tmp_path / "full.txt")
    assert full["template_directions"] == len(d["dir_rot"]) and m["template_directions"] <= len(d["dir_rot"])
    bank = json.loads((tmp_path / "bank.json").read_text())
    table = star.read_particle_table(str(tmp_path / "parts.star"))
    ctx = orientation_bank.settings(tmp_path / "parts.star", tmp_path / "dirs.star", 2, True, wl.local_range,
                                    wl.psi_step)
    orientation_bank.check_bank(bank, ctx)
    assert bank["rows"] == m["orientation_rows"] and set(bank["rows"]) == {str(r.line_index) for r in table.records}
    cal = _parse(tmp_path / "cal.txt")
    for line, row in bank["rows"].items():
        dirs = row["directions"]
        assert len(dirs) <= 2 and [x[3] for x in dirs] == sorted((x[3] for x in dirs), reverse=True)
        if int(line) in cal:  # the top-ranked direction is the record's best direction, its score the rawCC
            cc, raw = cal[int(line)]["cc: "], cal[int(line)]["rawCC: "]
            assert (dirs[0][1], dirs[0][2]) == (cc[5], cc[6]) and dirs[0][3] == raw[2]
    r = search.run_search(parse(base + ["--o", str(tmp_path / "res.txt"), "--orientations-in",
                                        str(tmp_path / "bank.json")]))
    assert r["orientation_rows"] is None
    assert r["template_directions"] == len({x[0] for row in bank["rows"].values() for x in row["directions"]})
    got = _parse(tmp_path / "res.txt")
    assert sorted(got) == sorted(cal)
    for line, g in got.items():
        c = cal[line]
        assert g["cc: "][5:7] == c["cc: "][5:7] and g["cc: "][8:10] == c["cc: "][8:10]
        assert g["psi: "] == c["psi: "] and g["rawCC: "] == c["rawCC: "]
    with pytest.raises(ValueError):
        search.run_search(parse(base + ["--o", str(tmp_path / "bad.txt")]), orientation_rows=bank["rows"],
                          calibrate=True)
    with pytest.raises(ValueError):  # a bank for other settings is refused
        search.run_search(parse(base[:-1] + ["3", "--o", str(tmp_path / "bad.txt"), "--orientations-in",
                                             str(tmp_path / "bank.json")]))


def test_pdb2mrc_cli_matches_generate_volume(hip_device, tmp_path):
    """main(--dontWriteMRC) returns generate_volume's density; with --o it
    writes the same density as an MRC."""
    from gisapr_amd import density, mrc, synthetic
    print("[case] pdb2mrc CLI", flush=True)
    d = synthetic.make_workload("box32")
    (tmp_path / "m.pdb").write_text(synthetic.pdb_text(d["coords"], d["elements"], d["chains"]))
    argv = ["--i", str(tmp_path / "m.pdb"), "--o", str(tmp_path / "m.mrc"), "--box", "32", "--apix", "3.0",
            "--res", "6.0"]
    arr = density.main(argv + ["--dontWriteMRC"])
    want = density.generate_volume(str(tmp_path / "m.pdb"), 32, 3.0, 6.0, device=str(hip_device)).cpu().numpy()
    np.testing.assert_allclose(arr, want, rtol=0, atol=1e-6 * np.abs(want).max())
    assert density.main(argv) == 0
    np.testing.assert_allclose(mrc.read(tmp_path / "m.mrc"), want, rtol=0, atol=1e-6 * np.abs(want).max())
