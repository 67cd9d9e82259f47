"""The reference's HDF5 container end to end on the GPU: a run written by OutputManager
(format "h5") loads back through dpmd_h5_load (BaseParticle::load_from_hdf5) into a fresh
context with every field bit-identical, and the loaded state then integrates bit for bit like
the oracle built from the same file."""
import numpy as np
import pytest

import oracle as orc
from dpmd_amd import _lib as L
from dpmd_amd import h5
from dpmd_amd import workloads as W
from dpmd_amd.output import OutputManager
from dpmd_amd.system import Disk, RigidBumpy, VelocityVerlet, run_fire

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not h5.available(), reason="libdpmd_h5 not built")]
C = L.ORDER_CANONICAL


def write_run(P, path, steps=60):
    om = OutputManager(P, path, max_workers=2, format="h5")
    om.set_trajectory_fields(["pos", "vel", "pe_total"])
    om.set_trajectory_interval(10)
    om.set_restart_interval(25)
    om.set_extra_final_fields(["pe", "packing_fraction"])
    om.initialize()
    vv = VelocityVerlet(P, 1e-2)
    vv.init()
    for k in range(1, steps + 1):
        vv.step(1)
        om.step(k)
    om.finalize()


def disk_written(path):
    b = W.disk_batch([64, 200, 97], [0.84, 0.86, 0.83], [0, 0, 0], seed=21)
    P = Disk()
    W.load_into(P, b)
    P.init_neighbors()
    run_fire(P, 1e-2, 300, L.FIRE_PRED_MAX_FORCE, 1e-12, mode=L.FORCE_PBC)
    rng = np.random.default_rng(2)
    P.set("VEL_X", rng.standard_normal(b.N) * 0.02)
    P.set("VEL_Y", rng.standard_normal(b.N) * 0.02)
    print(f"disk N={b.N} S={b.S}", flush=True)
    write_run(P, path)
    return b, P


def test_disk_h5_load_matches_written_state_and_oracle(tmp_path):
    b, P = disk_written(tmp_path)
    run = h5.read_run(tmp_path)
    assert run["static"]["class_name"] == "Disk" and run["static"]["neighbor_method"] == "Cell"
    # finalize() writes a last restart snapshot at the final step, as the reference does
    assert run["restart"]["step"] == 60 and np.array_equal(run["timestep"], np.arange(10, 61, 10))
    for loc in ("init", "restart", "final"):
        Q = Disk()
        h5.load_from_hdf5(Q, tmp_path, loc)
        assert Q.neighbor_method == L.NEIGHBOR_CELL and Q.n_particles() == b.N and Q.n_systems() == b.S
        g = run[loc]
        assert np.array_equal(Q.get("POS_X", C), g["pos"][:, 0]) and np.array_equal(Q.get("VEL_Y", C), g["vel"][:, 1])
        assert np.array_equal(Q.get("RAD", C), run["static"]["rad"])
        assert np.array_equal(Q.get("BOX_SIZE_X"), run["static"]["box_size"][:, 0])
        assert np.array_equal(Q.get("SYSTEM_OFFSET"), run["static"]["system_offset"])
    # Q holds the final state; it equals the writer's state and continues like the oracle built
    # from the same file
    for f in ("POS_X", "POS_Y", "VEL_X", "VEL_Y", "FORCE_X"):
        assert np.array_equal(Q.get(f, C), P.get(f, C)), f
    s, fin = run["static"], run["final"]
    cell = s["cell_dim"]
    b2 = W.DiskBatch(off=s["system_offset"], rad=s["rad"], mass=s["mass"], area=s["area"],
                     Lx=s["box_size"][:, 0], Ly=s["box_size"][:, 1], e=s["e_interaction"],
                     cutoff=s["neighbor_cutoff"], thresh2=s["thresh2"], boundary=np.zeros(b.S, np.int32),
                     x=fin["pos"][:, 0], y=fin["pos"][:, 1], cdx=cell[:, 0], cdy=cell[:, 1])
    ref = orc.OracleDisk(b2)
    assert np.array_equal(ref.static_index(), Q.get("STATIC_INDEX"))
    ref.set("vx", Q.get("VEL_X"))
    ref.set("vy", Q.get("VEL_Y"))
    VelocityVerlet(Q, 1e-2).init()
    ref.vv_init(1e-2, L.FORCE_PBC)
    VelocityVerlet(Q, 1e-2).run(200)
    ref.vv_run(200, L.FORCE_PBC)
    for f, r in (("POS_X", "x"), ("POS_Y", "y"), ("VEL_X", "vx"), ("VEL_Y", "vy")):
        assert np.array_equal(Q.get(f), ref.get(r)), f


def test_disk_h5_load_rejects_inconsistent_files(tmp_path):
    disk_written(tmp_path)
    with h5.H5File(tmp_path / "meta.h5", h5.READ_WRITE) as f:
        sid = f.read("/static/system_id")
        sid[5] += 1
        f.write("/static/system_id", sid)
    with pytest.raises(ValueError, match="system_id disagrees"):
        h5.load_from_hdf5(Disk(), tmp_path, "final")
    with pytest.raises(ValueError, match="invalid location"):
        h5.load_from_hdf5(Disk(), tmp_path, "middle")
    with pytest.raises(ValueError, match="holds a Disk, not a RigidBumpy"):
        h5.load_from_hdf5(RigidBumpy(), tmp_path, "final")
    with pytest.raises(ValueError, match="does not exist"):
        h5.load_from_hdf5(Disk(), tmp_path / "nowhere", "init")


def test_rigid_h5_load_matches_written_state_and_oracle(tmp_path):
    rb = W.rigid_batch(S=2, n_per=16, seed=5, boundary=0, phi=0.70)
    P = RigidBumpy()
    W.load_rigid_into(P, rb)
    P.init_neighbors()
    rng = np.random.default_rng(9)
    for f, sc in (("VEL_X", 0.03), ("VEL_Y", 0.03), ("ANGULAR_VEL", 0.05)):
        P.set(f, rng.standard_normal(rb.N) * sc)
    print(f"rigid N={rb.N} Nv={rb.Nv}", flush=True)
    write_run(P, tmp_path, steps=40)
    run = h5.read_run(tmp_path)
    assert run["static"]["class_name"] == "RigidBumpy" and run["static"]["n_vertices"] == rb.Nv
    Q = RigidBumpy()
    h5.load_from_hdf5(Q, tmp_path, "final")
    for f in ("POS_X", "POS_Y", "ANGLE", "VEL_X", "ANGULAR_VEL", "VERTEX_POS_X", "VERTEX_POS_Y", "VERTEX_RAD",
              "MOMENT_INERTIA", "N_VERTICES_PER_PARTICLE", "VERTEX_PARTICLE_ID"):
        assert np.array_equal(Q.get(f, C), P.get(f, C)), f
    s, fin = run["static"], run["final"]
    cell = s["cell_dim"]
    rb2 = W.RigidBatch(off=s["system_offset"], nvpp=s["n_vertices_per_particle"], x=fin["pos"][:, 0],
                       y=fin["pos"][:, 1], angle=fin["angle"], mass=s["mass"], inertia=s["moment_inertia"],
                       area=s["area"], vx=fin["vertex_pos"][:, 0], vy=fin["vertex_pos"][:, 1], vrad=s["vertex_rad"],
                       vmass=s["vertex_mass"], Lx=s["box_size"][:, 0], Ly=s["box_size"][:, 1], e=s["e_interaction"],
                       cutoff=s["neighbor_cutoff"], thresh2=s["thresh2"], boundary=np.zeros(rb.S, np.int32),
                       cdx=cell[:, 0], cdy=cell[:, 1])
    ref = orc.OracleRigid(rb2)
    for f, r in (("VEL_X", "vx"), ("VEL_Y", "vy"), ("ANGULAR_VEL", "w")):
        ref.set(r, Q.get(f))
    VelocityVerlet(Q, 1e-2).init()
    ref.vv_init(1e-2, L.FORCE_PBC)
    VelocityVerlet(Q, 1e-2).run(150)
    ref.vv_run(150, L.FORCE_PBC)
    for f, r in (("POS_X", "x"), ("ANGLE", "angle"), ("VEL_Y", "vy"), ("ANGULAR_VEL", "w"), ("VERTEX_POS_X", "vpx")):
        assert np.array_equal(Q.get(f), ref.get(r)), f
