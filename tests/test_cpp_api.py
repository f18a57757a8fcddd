"""The C++ mirror of the reference API (include/reactphysics3d_mi355x.hpp): the helloworld example
builds against it, constructs its world through the C-ABI, and (GPU) steps it like the reference."""
import os
import subprocess

import numpy as np
import pytest

from reactphysics3d_amd import scene as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "reactphysics3d_amd", "_lib", "hello_world")


def _run(steps):
    if not os.path.exists(EXE):
        pytest.skip("example not built (python -m reactphysics3d_amd.build)")
    out = subprocess.run([EXE, str(steps)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    rows = [list(map(float, line.split()[2:])) for line in out.stdout.splitlines() if line.startswith("body")]
    return np.array(rows, np.float32)


def _scene():
    # WorldSettings defaults: sleeping on
    sc = S.Scene("hello_world", S.make_settings(gravity=(0.0, 0.0, -9.81), vel_iters=10, pos_iters=5, sleeping=1))
    g = sc.add_body(S.STATIC, (0.0, 0.0, -0.5))
    sc.add_collider(g, S.BOX, (10.0, 10.0, 0.5), friction=0.5, bounciness=0.0)
    b = sc.add_body(S.DYNAMIC, (0.0, 0.0, 2.0), (0.0998334, 0.0, 0.0, 0.9950042))
    sc.add_collider(b, S.BOX, (0.5, 0.5, 0.5), friction=0.5, bounciness=0.0, density=1.0)
    s = sc.add_body(S.DYNAMIC, (3.5, 0.0, 3.0), lin=(-1.0, 0.0, 0.0))
    sc.add_collider(s, S.SPHERE, (0.4,), friction=0.5, bounciness=0.0, density=1.0)
    return sc


def test_example_builds_and_creates_world():
    st = _run(0)
    assert st.shape == (3, 13)
    assert np.allclose(st[1, :7], [0, 0, 2, 0.0998334, 0, 0, 0.9950042], atol=1e-7)
    assert np.allclose(st[2, 7:10], [-1, 0, 0])


@pytest.mark.gpu
def test_example_steps_like_the_reference(ref_available):
    from oracle.ref import RefWorld
    st = _run(120)
    ref = RefWorld(_scene())
    for _ in range(120):
        ref.step(1.0 / 60.0)   # box and ball never touch: their manifolds share only the static ground
    p, q, v, w = ref.bodies()
    want = np.concatenate([p, q, v, w], 1)
    err = np.max(np.abs(st - want))
    print("max |device - reference|", err)
    assert err <= 1e-4


JOINTS_EXE = os.path.join(ROOT, "reactphysics3d_amd", "_lib", "joints")


def _run_joints(steps, baumgarte=False):
    if not os.path.exists(JOINTS_EXE):
        pytest.skip("example not built (python -m reactphysics3d_amd.build)")
    out = subprocess.run([JOINTS_EXE, str(steps), "1" if baumgarte else "0"], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr
    rows = [list(map(float, line.split()[2:])) for line in out.stdout.splitlines() if line.startswith("body")]
    return np.array(rows, np.float32)


def _joints_scene(baumgarte=False):
    """The world of examples/joints.cpp as a scene description."""
    bg = S.JOINT_BAUMGARTE if baumgarte else 0
    sc = S.Scene("joints_example", S.make_settings(gravity=(0.0, 0.0, -9.81), vel_iters=10, pos_iters=5, sleeping=1))
    post = sc.add_body(S.STATIC, (0.0, 0.0, 5.0))
    car = sc.add_body(S.DYNAMIC, (0.0, 0.0, 4.2))
    sc.add_collider(car, S.BOX, (0.3, 0.3, 0.3))
    pivot = sc.add_body(S.STATIC, (4.0, 0.0, 5.0))
    paddle = sc.add_body(S.DYNAMIC, (4.6, 0.0, 5.0))
    sc.add_collider(paddle, S.BOX, (0.5, 0.1, 0.3))
    hook = sc.add_body(S.STATIC, (8.0, 0.0, 5.0))
    ball = sc.add_body(S.DYNAMIC, (9.0, 0.0, 5.0))
    sc.add_collider(ball, S.SPHERE, (0.2,))
    sc.add_joint(S.SLIDER, post, car, (0.0, 0.0, 4.2), axis=(0.6, 0.0, 0.8), flags=3 | bg, limits=(-0.4, 0.6, 0.5, 1.0))
    sc.add_joint(S.HINGE, pivot, paddle, (4.05, 0.0, 5.0), axis=(0.0, 1.0, 0.0), flags=1 | bg,
                 limits=(-0.6, 0.4, 0.0, 0.0))
    sc.add_joint(S.BALLSOCKET, hook, ball, (8.5, 0.0, 5.0), flags=1 | bg, limits=(0.3, 0.0, 0.0, 0.0))
    return sc


def test_joints_example_builds_and_creates_world():
    st = _run_joints(0)
    assert st.shape == (6, 13)
    assert np.allclose(st[:, 2], 5.0 - 0.8 * (np.arange(6) == 1))


@pytest.mark.gpu
@pytest.mark.parametrize("baumgarte", [False, True])
def test_joints_example_steps_like_the_reference(ref_available, baumgarte):
    """Limits and motors set through SliderJointInfo fields and through the joint setters after
    createJoint reach the device, and so does JointInfo::positionCorrectionTechnique: the example matches
    the reference world built with the same limits (over 240 steps, long enough for the pendulums at rest
    on their limits to fall asleep)."""
    from oracle.ref import RefWorld
    st = _run_joints(240, baumgarte)
    ref = RefWorld(_joints_scene(baumgarte))
    for _ in range(240):
        ref.step(1.0 / 60.0)
    print("asleep in the reference:", ref.sleeping())
    p, q, v, w = ref.bodies()
    want = np.concatenate([p, q, v, w], 1)
    err = np.max(np.abs(st - want))
    print("max |device - reference|", err)
    assert err <= 1e-4


CONTACTS_EXE = os.path.join(ROOT, "reactphysics3d_amd", "_lib", "contacts")


def _contacts_scene():
    """The world of examples/contacts.cpp (WorldSettings defaults, gravity -z).  The boxes never touch
    each other, so no two manifolds share a dynamic body and the reference's own solve order gives
    the device's results."""
    sc = S.Scene("contacts_example", S.make_settings(gravity=(0.0, 0.0, -9.81), sleeping=1))
    g = sc.add_body(S.STATIC, (0.0, 0.0, -0.5))
    sc.add_collider(g, S.BOX, (10.0, 10.0, 0.5))
    for p in ((0.0, 0.0, 1.0), (3.0, 0.0, 2.5), (-3.0, 0.5, 4.0)):
        b = sc.add_body(S.DYNAMIC, p)
        sc.add_collider(b, S.BOX, (0.5, 0.5, 0.5))
    return sc


@pytest.mark.gpu
def test_event_listener_sees_the_reference_events(ref_available):
    """An EventListener passed to PhysicsWorld::setEventListener of the C++ mirror receives, every
    step, the reference's counts of ContactStart / ContactStay / ContactExit pairs and contact points."""
    from oracle.ref import RefWorld
    if not os.path.exists(CONTACTS_EXE):
        pytest.skip("example not built (python -m reactphysics3d_amd.build)")
    out = subprocess.run([CONTACTS_EXE, "150"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    rows = np.array([list(map(int, line.split()[2:])) for line in out.stdout.splitlines() if line.startswith("step")])
    ref = RefWorld(_contacts_scene())
    ref.enable_events()
    want = []
    for _ in range(150):
        ref.step(1.0 / 60.0)
        ev = ref.events()
        want.append([int((ev[:, 2] == t).sum()) for t in range(3)] + [int(ev[:, 3].sum())])
    want = np.array(want)
    assert np.array_equal(rows, want)
    # PhysicsWorld::raycast through the mirror: the closest hit straight down onto the first box
    ray = [float(line.split()[1]) for line in out.stdout.splitlines() if line.startswith("ray")]
    col, pt, nr, fr = ref.raycast(np.array([[0, 0, 10]], np.float32), np.array([[0, 0, -10]], np.float32))
    assert col[0] == 1 and len(ray) == 1 and np.float32(ray[0]) == fr[0]
    assert want[:, 0].sum() >= 4 and want[:, 1].sum() > 100 and want[:, 2].sum() >= 1


DRIVEN_EXE = os.path.join(ROOT, "reactphysics3d_amd", "_lib", "driven")


def _driven_scene():
    """The world of examples/driven.cpp (WorldSettings defaults apart from gravity -z and 10 / 5
    iterations; the box and the ball never touch, so no two manifolds share a dynamic body)."""
    sc = S.Scene("driven_example", S.make_settings(gravity=(0.0, 0.0, -9.81), vel_iters=10, pos_iters=5, sleeping=1))
    g = sc.add_body(S.STATIC, (0.0, 0.0, -0.5))
    sc.add_collider(g, S.BOX, (20.0, 20.0, 0.5), friction=0.5, bounciness=0.0)
    b = sc.add_body(S.DYNAMIC, (0.0, 0.0, 0.6))
    sc.add_collider(b, S.BOX, (0.5, 0.5, 0.5), friction=0.5, bounciness=0.0)
    s = sc.add_body(S.DYNAMIC, (6.0, 0.0, 0.5))
    sc.add_collider(s, S.SPHERE, (0.4,), friction=0.5, bounciness=0.0)
    return sc


def _driven_calls(i):
    """The calls examples/driven.cpp makes before step i, as (code, body, flags, a, c)."""
    from reactphysics3d_amd import lib as L
    f32 = lambda *v: np.array(v, np.float32)
    ops = []
    if 20 <= i < 40:
        ops.append((L.OP_APPLY_FORCE, 1, L.FORCE_AT_POINT | L.FORCE_POINT_LOCAL | L.FORCE_LOCAL, f32(12, 0, 0), f32(0, 0.3, 0.4)))
    if i == 30:
        ops.append((L.OP_APPLY_TORQUE, 1, 0, f32(0, 0, 3), None))
    if i == 35:
        ops.append((L.OP_APPLY_FORCE, 1, L.FORCE_AT_POINT, f32(0, 4, 0), f32(0, 0, 1)))
    if i == 38:
        ops.append((L.OP_RESET_TORQUE, 1, 0, None, None))
    if i == 25:
        ops.append((L.OP_SET_LINEAR_VELOCITY, 2, 0, f32(-1, 2, 0), None))
    if i == 50:
        ops.append((L.OP_SET_TRANSFORM, 2, 0, f32(4, 4, 2), f32(0, 0.6, 0, 0.8)))
    if i == 51:
        ops.append((L.OP_SET_ANGULAR_VELOCITY, 2, 0, f32(0, 5, 0), None))
    if i == 52:
        ops.append((L.OP_APPLY_FORCE, 2, L.FORCE_LOCAL, f32(0, 0, 30), None))
    if i == 53:
        ops.append((L.OP_APPLY_TORQUE, 2, L.FORCE_LOCAL, f32(1, 0, 0), None))
    if i == 60:
        ops.append((L.OP_SET_SLEEPING, 1, 1, None, None))
    if i == 90:
        ops.append((L.OP_APPLY_FORCE, 1, 0, f32(0, 0, 20), None))
    if i == 91:
        ops.append((L.OP_APPLY_FORCE, 1, L.FORCE_AT_POINT | L.FORCE_LOCAL, f32(5, 0, 0), f32(0, 0, 0)))
    if i == 92:
        ops.append((L.OP_RESET_FORCE, 1, 0, None, None))
    return ops


@pytest.mark.gpu
def test_driven_example_calls_like_the_reference(ref_available):
    """RigidBody::setTransform / set*Velocity / apply*Force* / apply*Torque / reset* / setIsSleeping made
    between update() calls through the C++ mirror: the force and torque sums after the calls of a
    frame, the final states and the sleeping flags match the reference making the same calls."""
    from oracle.ref import RefWorld
    if not os.path.exists(DRIVEN_EXE):
        pytest.skip("example not built (python -m reactphysics3d_amd.build)")
    steps = 150
    probes = (20, 35, 38, 53, 91)
    ref = RefWorld(_driven_scene())
    want_probe = {}
    for i in range(steps):
        ref.body_ops(_driven_calls(i))
        if i in probes:
            f, t = ref.force_torque()
            want_probe[i] = np.concatenate([f[1], t[1]])
        if i == 61:
            assert ref.sleeping()[1] == 1
        ref.step(1.0 / 60.0)
    p, q, v, w = ref.bodies()
    want = np.concatenate([p, q, v, w], 1)
    for k, probe in enumerate(probes):
        out = subprocess.run([DRIVEN_EXE, str(steps), str(probe)], capture_output=True, text=True, timeout=120)
        assert out.returncode == 0, out.stderr
        lines = out.stdout.splitlines()
        got = np.array([list(map(float, line.split()[2:])) for line in lines if line.startswith("probe")][0], np.float32)
        assert np.array_equal(got, want_probe[probe]), f"frame {probe}: {got} vs {want_probe[probe]}"
        if k == 0:
            st = np.array([list(map(float, line.split()[2:])) for line in lines if line.startswith("body")], np.float32)
            err = np.max(np.abs(st - want))
            print("max |device - reference|", err)
            assert err <= 1e-4
            sl = [list(map(int, line.split()[1:])) for line in lines if line.startswith("sleeping")][0]
            assert sl == list(ref.sleeping()[1:])


MESHES_EXE = os.path.join(ROOT, "reactphysics3d_amd", "_lib", "meshes")


def test_convex_mesh_signatures_of_the_reference():
    """PolygonVertexArray with strided double vertices and short indices, VertexArray + QuickHull and
    createConvexMesh(..., messages) as the reference declares them (PolygonVertexArray.h:102,
    VertexArray.h:66, PhysicsCommon.h:230-233): the box mesh is accepted, the hull's polygons are the
    library's QuickHull polygons of the same points, the open mesh is refused with one error message."""
    from reactphysics3d_amd import lib
    if not os.path.exists(MESHES_EXE):
        pytest.skip("example not built (python -m reactphysics3d_amd.build)")
    out = subprocess.run([MESHES_EXE], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    lines = out.stdout.splitlines()
    get = lambda tag: [l.split()[1:] for l in lines if l.startswith(tag + " ")]
    assert get("box") == [["1", "8", "6"]]
    corners = [[1.0 if i & 1 else -1.0, 1.0 if i & 2 else -1.0, 1.0 if i & 4 else -1.0] for i in range(8)]
    extra = [[0, 0, 0], [0.5, 0.25, 1], [-0.25, 1, 0.5], [1, 0, 0], [0.3, -1, 0.1], [0, 0, -1]]
    v, c, idx = lib.quickhull(np.array(corners + extra, np.float32))
    assert get("hull") == [["1", str(len(v)), str(len(c))]]
    want, o = [], 0
    for f, n in enumerate(c):
        want.append(v[idx[o:o + n]].ravel())
        o += n
    faces = get("face")
    assert len(faces) == len(want)
    for f, row in enumerate(faces):
        assert np.array_equal(np.array(row[1:], np.float32), want[f]), f"face {f}"
    assert get("open") == [["0", "1"]]
    assert get("message")[0][0] == "1"   # Message::Type::Error
    assert get("colliders") == [["0", "1"]]


QUERIES_EXE = os.path.join(ROOT, "reactphysics3d_amd", "_lib", "queries")


def _queries_scene():
    """The world of examples/queries.cpp (WorldSettings defaults, gravity -z): three boxes apart on a
    ground, so no two manifolds share a dynamic body."""
    sc = S.Scene("queries_example", S.make_settings(gravity=(0.0, 0.0, -9.81), sleeping=1))
    g = sc.add_body(S.STATIC, (0.0, 0.0, -0.5))
    sc.add_collider(g, S.BOX, (10.0, 10.0, 0.5))
    for i in range(3):
        b = sc.add_body(S.DYNAMIC, (3.0 * i - 3.0, 0.0, 1.0 + i))
        sc.add_collider(b, S.BOX, (0.5, 0.5, 0.5))
    return sc


@pytest.mark.gpu
def test_world_queries_through_the_cpp_mirror(ref_available):
    """PhysicsWorld::testOverlap / testCollision / raycast (closest-hit and every-hit callbacks) of the C++
    mirror between update() calls answer as the reference does, and the steps after them agree."""
    from oracle.ref import RefWorld
    if not os.path.exists(QUERIES_EXE):
        pytest.skip("example not built (python -m reactphysics3d_amd.build)")
    out = subprocess.run([QUERIES_EXE, "60", "30"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    lines = out.stdout.splitlines()
    rows = lambda tag: [l.split()[1:] for l in lines if l.startswith(tag + " ")]
    ref = RefWorld(_queries_scene())
    for _ in range(60):
        ref.step(1.0 / 60.0)
    want_touch = [int(ref.test_overlap_bodies(1, 0)), int(ref.test_overlap_bodies(1, 2))]
    assert [int(x) for x in rows("touching")[0]] == want_touch
    o1 = ref.test_overlap(2)
    o2 = ref.test_overlap(-1)
    want_overlap = sorted({(int(a), int(b), int(e)) for a, b, e in zip(*o1)}) + \
        sorted({(int(a), int(b), int(e)) for a, b, e in zip(*o2)})
    got = [tuple(int(x) for x in r) for r in rows("overlap")]
    n1 = len(o1[0])
    assert sorted(got[:n1]) + sorted(got[n1:]) == want_overlap
    c_all = ref.test_collision()
    c_pair = ref.test_collision(3, 0)
    got_c = rows("contact")
    assert len(got_c) == len(c_all["colliders"]) + len(c_pair["colliders"])
    want_c = {}
    for ct in (c_all, c_pair):
        for m in range(len(ct["colliders"])):
            n = int(ct["nb_points"][m])
            want_c.setdefault((int(ct["colliders"][m, 0]), int(ct["colliders"][m, 1]), int(ct["event"][m]), n), []).append(
                ct["normal_depth"][m, :n].ravel())
    for r in got_c:
        key = tuple(int(x) for x in r[:4])
        assert key in want_c, f"unexpected contact pair {key}"
        vals = np.array(r[4:], np.float32)
        assert any(np.allclose(vals, w, atol=1e-5) for w in want_c[key])
    p1 = np.array([[-8.0, 0.0, 0.3]], np.float32)
    p2 = np.array([[8.0, 0.0, 0.3]], np.float32)
    cc, _, _, cf = ref.raycast(p1, p2)
    got_closest = rows("closest")
    assert len(got_closest) == 1 and int(got_closest[0][0]) == int(cc[0]) and np.float32(got_closest[0][1]) == cf[0]
    ray, col, _, _, fr = ref.raycast_all(p1, p2)
    got_hits = sorted((int(r[0]), float(np.float32(r[1]))) for r in rows("hit"))
    assert got_hits == sorted((int(c), float(f)) for c, f in zip(col, fr))
    for _ in range(30):
        ref.step(1.0 / 60.0)
    p, q, v, w = ref.bodies()
    st = np.array([list(map(float, r[1:])) for r in rows("body")], np.float32)
    err = np.max(np.abs(st - np.concatenate([p, q, v, w], 1)))
    print("max |device - reference|", err)
    assert err <= 1e-4


CONCAVE_EXE = os.path.join(ROOT, "reactphysics3d_amd", "_lib", "concave")


def _concave_scene():
    """The world of examples/concave.cpp (WorldSettings defaults, gravity -z)."""
    f = np.float32
    sc = S.Scene("concave_example", S.make_settings(gravity=(0.0, 0.0, -9.81), sleeping=1))
    n = 8
    v = np.array([(f(i - n // 2), f(j - n // 2), f(0.1) * f((7 * i + 3 * j) % 5) - f(0.2))
                  for j in range(n + 1) for i in range(n + 1)], np.float32)
    t = []
    for j in range(n):
        for i in range(n):
            a = j * (n + 1) + i
            b, c, d = a + 1, a + n + 2, a + n + 1
            t += [(a, b, c), (a, c, d)] if (i + j) % 2 == 0 else [(a, b, d), (b, c, d)]
    m = sc.add_triangle_mesh(v, np.array(t, np.int32))
    g = sc.add_body(S.STATIC, (0.0, 0.0, 0.0))
    sc.add_collider(g, S.CONCAVE_MESH, (float(m), 1.0, 1.0, 1.0))
    h = np.array([[f(0.05) * f((x + 2 * y) % 4) for x in range(9)] for y in range(9)], np.float32)
    hf = sc.add_heightfield(h)
    hb = sc.add_body(S.STATIC, (12.0, 0.0, 0.0))
    sc.add_collider(hb, S.HEIGHTFIELD, (float(hf), 1.0, 1.5, 1.0), lquat=(0.70710677, 0.0, 0.0, 0.70710677))
    drops = [(-2, -1, 2), (1, 2, 3), (2, -2, 2.5), (-1, 1.5, 4), (11, 1, 2), (13, -1, 3)]
    for k, p in enumerate(drops):
        b = sc.add_body(S.DYNAMIC, tuple(float(x) for x in p))
        if k < 2:
            sc.add_collider(b, S.SPHERE, (0.4,))
        elif k == 3:
            sc.add_collider(b, S.CAPSULE, (0.25, 0.8))
        else:
            sc.add_collider(b, S.BOX, (0.4, 0.3, 0.35))
    return sc


def test_concave_example_builds_and_creates_world():
    if not os.path.exists(CONCAVE_EXE):
        pytest.skip("example not built (python -m reactphysics3d_amd.build)")
    out = subprocess.run([CONCAVE_EXE, "0"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    rows = [list(map(float, line.split()[2:])) for line in out.stdout.splitlines() if line.startswith("body")]
    assert len(rows) == 8 and rows[2][:3] == [-2.0, -1.0, 2.0]


@pytest.mark.gpu
def test_concave_example_steps_like_the_reference(ref_available):
    """TriangleVertexArray / createTriangleMesh / createConcaveMeshShape and createHeightField /
    createHeightFieldShape through the C++ mirror: the example matches the reference world built with the
    same data over 150 steps."""
    from oracle.ref import RefWorld
    out = subprocess.run([CONCAVE_EXE, "150"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    st = np.array([list(map(float, line.split()[2:])) for line in out.stdout.splitlines() if line.startswith("body")],
                  np.float32)
    ref = RefWorld(_concave_scene())
    for _ in range(150):
        ref.step(1.0 / 60.0)
    p, q, v, w = ref.bodies()
    want = np.concatenate([p, q, v, w], 1)
    err = np.max(np.abs(st - want))
    print("max |device - reference|", err, "resting heights", want[2:, 2])
    assert want[2:, 2].max() < 1.0   # everything landed
    assert err <= 1e-4
