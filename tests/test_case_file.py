"""The case runner as the caller of the path: the reference's case file
(`cases/tidal_channel/case.toml`, a committed fixture) read into the same
problem as `cases.tidal_channel`, `--set` overrides, refusals by name; the
runner's station sampling (point location, Newton inverse map, the nodal basis
at a point) and its `stations.csv` number format; the `DGRST2D` restart file
(published FNV-1a test vectors, round trip, damage and domain checks); the
runner on the restatement: 5,765 steps and 51 records per station."""
import math
import os

import numpy as np
import pytest

from dgcom_amd import case_file as CF, cases, restart as R
import oracle as O

CASE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "tidal_channel_case.toml")


class _HostState:
    def __init__(self, x):
        self.x = x

    def download(self):
        return tuple(a.copy() for a in self.x)

    def is_finite(self):
        return all(np.isfinite(a).all() for a in self.x)

    def free(self):
        pass


class OraclePhysics:
    """The restatement behind the runner's physics calls (state, dt, step)."""

    def __init__(self, case):
        self.orc = O.Oracle.from_case(case)

    def state(self, h, hu, hv):
        return _HostState([np.array(h, dtype=np.float64), np.array(hu, dtype=np.float64),
                           np.array(hv, dtype=np.float64)])

    def compute_dt_ssp(self, s, cfl, scheme):
        return self.orc.dt(*s.x, cfl, scheme)

    def step(self, s, t, dt, scheme):
        self.orc.run_dts(*s.x, t, [dt], scheme)

    def negative_depth_clips(self):
        return self.orc.clips()


def test_the_case_file_is_the_tidal_channel():
    run = CF.load(CASE)
    c, d = run.case, cases.tidal_channel()
    assert (run.name, run.duration_s, run.interval_s) == ("tidal_channel", 90_000.0, 1800.0)
    assert [s.name for s in run.stations] == ["mouth", "middle", "head"]
    assert [s.open_water_depth for s in run.stations] == [3.0, 3.0, 1.0]
    for a in ("bed", "h", "hu", "hv", "manning"):
        assert np.array_equal(getattr(c, a), getattr(d, a)), a
    for a in ("corners", "links", "tags"):
        assert np.array_equal(getattr(c.mesh, a), getattr(d.mesh, a)), a
    assert (c.f0, c.cfl, c.scheme, c.order, c.incoming_wave) == (d.f0, d.cfl, d.scheme, d.order, False)
    assert c.tide == d.tide


def test_overrides_and_refusals():
    run = CF.load(CASE, ["boundary.0.velocity=incoming_wave", "numerics.local_time_stepping_levels=3",
                         "run.duration_hours=1.5", "numerics.integrator=ssp_rk3"])
    assert run.case.incoming_wave and run.lts_levels == 3 and run.duration_s == 5400.0
    assert run.case.scheme == 0 and run.case.cfl == 0.9 * 1.04
    with pytest.raises(CF.CaseError, match="wind"):
        CF.load(CASE, ["physics.wind=1.0"])
    with pytest.raises(CF.CaseError, match="waves"):
        CF.load(CASE, ["waves.model=1"])
    with pytest.raises(CF.CaseError, match="drop_land"):
        CF.load(CASE, ["mesh.drop_land=true"])


def test_rust_number_format():
    # Rust's `{}` for f64: shortest round-trip digits, never an exponent
    for x, s in [(5.0, "5"), (0.1, "0.1"), (1e-7, "0.0000001"), (-0.0, "-0"), (1e21, "1000000000000000000000"),
                 (1800.0, "1800"), (0.30000000000000004, "0.30000000000000004"), (2.5e-5, "0.000025"),
                 (-1.25, "-1.25"), (123456789.125, "123456789.125"), (float("nan"), "NaN"), (float("inf"), "inf")]:
        assert CF.rust_display(x) == s, (x, CF.rust_display(x), s)
    for x in np.random.default_rng(1).standard_normal(200) * 10.0 ** np.random.default_rng(2).integers(-12, 12, 200):
        assert float(CF.rust_display(float(x))) == float(x)


def test_points_and_the_nodal_basis():
    # Newton's inverse map is exact in one step on a parallelogram, and the
    # nodal basis at a point reproduces a polynomial of degree <= N in r and s
    mesh = cases.structured_mesh(0.0, 3.0, 0.0, 2.0, 3, 2, jitter=0.2, seed=3)
    loc = CF.PointLocator(mesh)
    rng = np.random.default_rng(4)
    for order in (1, 2, 3):
        x, y = cases.node_coordinates(mesh, order)
        field = (1.0 + 0.5 * x - 0.25 * y).ravel()
        for _ in range(20):
            p = (rng.uniform(0.05, 2.95), rng.uniform(0.05, 1.95))
            k, r, s = loc.locate(p)
            cr = mesh.corners[k]
            # the point maps back to itself
            n = [(1 - r) * (1 - s) / 4, (1 + r) * (1 - s) / 4, (1 + r) * (1 + s) / 4, (1 - r) * (1 + s) / 4]
            assert abs(sum(n[i] * cr[i, 0] for i in range(4)) - p[0]) < 1e-12
            assert abs(sum(n[i] * cr[i, 1] for i in range(4)) - p[1]) < 1e-12
            pr = CF.Probe(k, r, s, order)
            assert abs(sum(pr.weights) - 1.0) < 1e-13
            nn = (order + 1) ** 2
            # x and y are bilinear in (r, s) on every quad: exact for order >= 1
            assert abs(pr.evaluate(field[k * nn:(k + 1) * nn]) - (1.0 + 0.5 * p[0] - 0.25 * p[1])) < 1e-12
    assert loc.locate((3.5, 1.0)) is None
    par = [(0.0, 0.0), (2.0, 0.5), (2.5, 2.0), (0.5, 1.5)]
    r, s = CF.inverse_bilinear(par, (1.25, 1.0))
    assert abs(r) < 1e-15 and abs(s) < 1e-15


def test_the_stations_sit_at_their_elements_corners():
    run = CF.load(CASE)
    probes = dict(CF.station_probes(run))
    assert [p.element for p in probes.values()] == [41, 59, 73]
    for p in probes.values():
        assert p.weights[-1] == 1.0 and all(w == 0.0 for w in p.weights[:-1])


def test_fnv1a_published_vectors():
    assert R.fnv1a(b"") == 0xCBF29CE484222325
    assert R.fnv1a(b"a") == 0xAF63DC4C8601EC8C
    assert R.fnv1a(b"foobar") == 0x85944171F73967E8
    assert R.fnv1a(b"bar", R.fnv1a(b"foo")) == R.fnv1a(b"foobar")


def test_restart_round_trip_and_checks(tmp_path):
    c = cases.drying_beach(nx=6, ny=3)
    rng = np.random.default_rng(5)
    h, hu, hv = (rng.standard_normal(c.h.size) for _ in range(3))
    fp = R.domain_fingerprint(c.mesh, c.bed)
    rst = R.Restart2D(123.25, fp, h, hu, hv, c.n_elements, c.nn, metadata=[("steps", "7"), ("case", "beach")],
                      text=[("case.resolved.toml", "a = 1\n")], extra=[("station_records", np.arange(10.0))])
    path = str(tmp_path / "beach.dgrst")
    rst.write(path)
    assert not os.path.exists(path + ".tmp")
    back = R.Restart2D.read(path)
    assert back.time == 123.25 and back.domain == fp
    assert np.array_equal(back.h, h) and np.array_equal(back.hu, hu) and np.array_equal(back.hv, hv)
    assert dict(back.metadata) == {"steps": "7", "case": "beach"}
    assert dict(back.text) == {"case.resolved.toml": "a = 1\n"}
    assert np.array_equal(dict(back.extra)["station_records"], np.arange(10.0))
    back.check_domain(c.mesh, c.bed, c.order)
    data = bytearray(open(path, "rb").read())
    assert data[:8] == b"DGRST2D\x01"
    assert R.fnv1a(bytes(data[:-8])) == int.from_bytes(data[-8:], "little")
    data[100] ^= 1
    open(path, "wb").write(bytes(data))
    with pytest.raises(R.RestartError, match="checksum"):
        R.Restart2D.read(path)
    bed2 = c.bed.copy()
    bed2[3] = np.nextafter(bed2[3], 1.0)
    with pytest.raises(R.RestartError, match="another domain"):
        rst.check_domain(c.mesh, bed2, c.order)


def _reference_layout_restart(run, path, rep):
    """A restart of `rep`'s end with the driver data laid out as the reference
    runner writes them (src/case/run/restart.rs:191-213), built here key by key."""
    c = run.case
    meta = [("build", "dgcom 0.1.0 (reference)")] + [(f"station.{j}", s.name) for j, s in enumerate(run.stations)]
    extra = [("output.counters", np.array([3.0, 1.25e9]))]
    for j, s in enumerate(run.stations):
        rows = np.array([r[1:] for r in rep.records if r[0] == s.name])
        for col, field in enumerate(("time_s", "eta_m", "u_m_s", "v_m_s", "depth_m")):
            extra.append((f"station.{j}.{field}", rows[:, col].copy()))
        for field in ("hs_m", "tp_s", "wave_from_deg"):
            extra.append((f"station.{j}.{field}", np.zeros(0)))
    R.Restart2D(rep.t, R.domain_fingerprint(c.mesh, c.bed), *rep.state, c.n_elements, c.nn, metadata=meta,
                text=[("case", "# dgcom 0.1.0 (reference)\n[run]\nname = \"tidal_channel\"\n")],
                extra=extra).write(path)


def test_a_reference_layout_restart_resumes(tmp_path):
    """A restart in the reference runner's layout (no `steps`, the records in
    `station.<j>.<field>`, `output.counters`) resumes on the restatement: the
    earlier records reach stations.csv and the run equals a straight one bit for
    bit; the runner's own restart has the same keys; without the counters, with
    other stations or more of them it is refused by name."""
    run = CF.load(CASE, ["run.duration_hours=2.0"])
    straight = CF.run_case(run, out_dir=str(tmp_path / "straight"), physics=OraclePhysics(run.case))
    first = CF.run_case(run, stop_at=3600.0, physics=OraclePhysics(run.case))
    path = str(tmp_path / "ref.restart")
    _reference_layout_restart(run, path, first)
    rest = CF.run_case(run, out_dir=str(tmp_path / "rest"), restart_from=path, physics=OraclePhysics(run.case))
    assert rest.records == straight.records and len(rest.records) == 5 * 3
    assert rest.volume[0] == 1.25e9
    for a, b in zip(rest.state, straight.state):
        assert np.array_equal(a, b)
    assert open(tmp_path / "rest" / "stations.csv").read() == open(tmp_path / "straight" / "stations.csv").read()
    # the runner's own restart: the reference's keys (and its step count)
    own = str(tmp_path / "own.restart")
    CF.run_case(run, stop_at=3600.0, restart_to=own, physics=OraclePhysics(run.case))
    back = R.Restart2D.read(own)
    ref = R.Restart2D.read(path)
    assert [k for k, _ in back.extra] == [k for k, _ in ref.extra]
    assert [k for k, _ in back.metadata if k != "steps"] == [k for k, _ in ref.metadata]
    assert [k for k, _ in back.text] == ["case"]
    assert dict(back.extra)["output.counters"][1] == straight.volume[0]
    assert CF.load(dict(back.text)["case"].split("\n", 1)[1]).doc == run.doc
    resumed_own = CF.run_case(run, restart_from=own, physics=OraclePhysics(run.case))
    assert resumed_own.records == straight.records and resumed_own.steps == straight.steps
    # refusals, by name
    for drop, match in (("output.counters", "no output.counters"), ("station.1.eta_m", "no station.1.eta_m")):
        bad = R.Restart2D.read(path)
        bad.extra = [(k, v) for k, v in bad.extra if k != drop]
        bad.write(path + ".bad")
        with pytest.raises(CF.CaseError, match=match):
            CF.run_case(run, restart_from=path + ".bad", physics=OraclePhysics(run.case))
    bad = R.Restart2D.read(path)
    bad.metadata = [(k, "elsewhere" if k == "station.2" else v) for k, v in bad.metadata]
    bad.write(path + ".bad")
    with pytest.raises(CF.CaseError, match="station 2 is head, the restart's elsewhere"):
        CF.run_case(run, restart_from=path + ".bad", physics=OraclePhysics(run.case))
    bad.metadata = [(k, v) for k, v in ref.metadata] + [("station.3", "extra")]
    bad.write(path + ".bad")
    with pytest.raises(CF.CaseError, match="more stations"):
        CF.run_case(run, restart_from=path + ".bad", physics=OraclePhysics(run.case))


def test_the_runner_on_the_restatement():
    """The runner over the restatement: 5,765 steps to 25 h, 51 records per
    station (the start and every 30 min), no clip."""
    run = CF.load(CASE)
    rep = CF.run_case(run, physics=OraclePhysics(run.case))
    assert rep.steps == 5765 and rep.t == 90_000.0 and rep.clips == 0
    assert len(rep.records) == 51 * 3
    times = sorted({r[1] for r in rep.records})
    assert times == [1800.0 * i for i in range(51)]
    assert all(math.isfinite(v) for r in rep.records for v in r[1:])
