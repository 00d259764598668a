"""The differentiation workload at its own size: 1,048,576 particles from the
32-component mixture, one of the 64 views at 512 x 512, 4 spp expected
stochastic marching (aarbf_diff_forward / aarbf_diff_backward), against the
FP64 oracle (oracle/adjoint.c):

  - the forward on a 32-pixel crop to 1e-4 absolute;
  - the adjoint of the loss L = sum over the crop of g . C (g seeded) to 1e-3
    relative L2 per gradient block (positions, radii, attributes, colormap,
    densitymap), and exactly zero for every particle no crop ray meets;
  - central finite differences of the FP64 oracle's L on 1,000 random
    parameters (radii, attributes, transfer-function texels, positions)
    against the GPU's gradient, to 1e-3 relative.

The oracle evaluates each crop pixel on the particles whose spheres can meet
the pixel's ray (distance from the ray's line below the radius, with a
margin for the finite-difference steps): every other particle covers no
march point of that pixel, so the pixel's value and derivatives are the same
as over the whole scene (checked on a smaller scene in
test_oracle_adjoint.py)."""
import concurrent.futures as cf

import numpy as np
import pytest

import aarbf
from aarbf import scenes
from aarbf.diff import _Render
from helpers import near_ray, pixel_ray
import oracle

pytestmark = pytest.mark.gpu

THREADS = 16
NPIX = 32
STEPS = {"pos": 1e-6, "radii": 1e-7, "attr": 1e-6, "colormap": 1e-6, "densitymap": 1e-6}


@pytest.fixture(scope="module")
def c4(gpu_ctx):
    import torch
    dw = scenes.diff_workload()
    W, H = dw.width, dw.height
    lf, la, up = dw.view(0)
    cam = aarbf.camera_basis(lf, la, up, scenes.fovy_scale(45.0), W, H)
    T = lambda x: torch.from_numpy(np.ascontiguousarray(x)).cuda()
    r = _Render(gpu_ctx, cam, W, H, dw.unit, dw.spp, 0, dw.rbf_radius, 1, 8, T(dw.xyzw[:, :3]), T(dw.xyzw[:, 3]),
                T(dw.radii), T(dw.colormap), T(dw.densitymap), T(dw.radiusmap))
    img = r.forward().cpu().numpy()
    # the crop: 32 pixels spread over the brighter half of the lit pixels
    val = img.sum(2)
    lit = np.argwhere(val > 0.05)
    assert len(lit) > 1000
    lit = lit[val[lit[:, 0], lit[:, 1]] >= np.median(val[lit[:, 0], lit[:, 1]])]
    pick = lit[np.linspace(0, len(lit) - 1, NPIX).astype(np.int64)]
    px = np.stack([pick[:, 1], pick[:, 0]], 1).astype(np.int32)
    g = np.random.default_rng(3).normal(size=(NPIX, 3))
    gimg = torch.zeros((H, W, 3), dtype=torch.float32, device="cuda")
    gimg[torch.from_numpy(pick[:, 0]).cuda(), torch.from_numpy(pick[:, 1]).cuda()] = T(g.astype(np.float32))
    gp, ga, gr, gcm, gdm, grm = (x.cpu().numpy() for x in r.backward(gimg))
    aabb = np.asarray(r.aabb, np.float32)
    r.close()
    gpu = {"pos": gp, "attr": ga, "radii": gr, "colormap": gcm, "densitymap": gdm[:, 0]}
    # per-pixel particle subsets and oracle objects (caller's particle order)
    pos = dw.xyzw[:, :3].astype(np.float64)
    subsets, objs = [], []
    for x, y in px:
        o, d, t0, t1 = pixel_ray(cam, W, H, x, y, aabb, dw.rbf_radius)
        idx = near_ray(pos, dw.radii.astype(np.float64), o, d, t0, t1)
        subsets.append(idx)
        objs.append(oracle.Diff(pos[idx].copy(), dw.xyzw[idx, 3].astype(np.float64), dw.radii[idx].astype(np.float64),
                                dw.colormap, dw.densitymap, dw.radiusmap, cam, W, H, dw.unit, dw.rbf_radius, aabb,
                                spp=dw.spp, frame_id=0))
    return dict(dw=dw, cam=cam, img=img, px=px, g=g, gpu=gpu, subsets=subsets, objs=objs, n=len(dw.xyzw))


def _oracle_adjoint(c):
    n = c["n"]
    tot = {"pos": np.zeros((n, 3)), "radii": np.zeros(n), "attr": np.zeros(n), "colormap": np.zeros((256, 4)),
           "densitymap": np.zeros(256)}

    def one(k):
        return c["objs"][k].adjoint(c["g"][k:k + 1], pixels=c["px"][k:k + 1])

    with cf.ThreadPoolExecutor(THREADS) as ex:
        res = list(ex.map(one, range(NPIX)))
    for k, a in enumerate(res):
        idx = c["subsets"][k]
        np.add.at(tot["pos"], idx, a["pos"])
        for key in ("radii", "attr"):
            np.add.at(tot[key], idx, a[key])
        tot["colormap"] += a["colormap"]
        tot["densitymap"] += a["densitymap"]
    return tot


def test_c4_forward_and_adjoint_match_oracle(c4):
    c = c4
    with cf.ThreadPoolExecutor(THREADS) as ex:
        fw = list(ex.map(lambda k: c["objs"][k].forward(pixels=c["px"][k:k + 1])[0], range(NPIX)))
    got = c["img"][c["px"][:, 1], c["px"][:, 0]]
    err = np.abs(got - np.array(fw)).max()
    print(f"\nforward on the crop: max abs error {err:.2e}; subsets {[len(s) for s in c['subsets']]}")
    assert err < 1e-4
    og = _oracle_adjoint(c)
    used = np.zeros(c["n"], bool)
    for s in c["subsets"]:
        used[s] = True
    for k in ("pos", "radii", "attr", "colormap", "densitymap"):
        a, b = np.asarray(c["gpu"][k], np.float64), og[k]
        rel = float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))
        print(f"{k}: relative L2 error {rel:.2e}")
        assert rel < 1e-3, k
    for k in ("pos", "radii", "attr"):
        assert not np.any(c["gpu"][k][~used]), f"{k}: gradient outside the particles the crop's rays meet"


def _params(c, og, rng):
    """1,000 parameters drawn among those with a nonzero oracle gradient:
    250 radii, 250 attributes, 250 texels (150 colormap, 100 densitymap),
    the rest positions: (kind, index, component)."""
    out = []
    for kind, count in (("radii", 250), ("attr", 250)):
        cand = np.nonzero(og[kind])[0]
        out += [(kind, int(i), 0) for i in rng.choice(cand, size=min(count, len(cand)), replace=False)]
    cm = np.argwhere(og["colormap"] != 0)
    dm = np.nonzero(og["densitymap"])[0]
    sel = cm[rng.choice(len(cm), size=min(150, len(cm)), replace=False)]
    out += [("colormap", int(i), int(j)) for i, j in sel]
    out += [("densitymap", int(i), 0) for i in rng.choice(dm, size=min(100, len(dm)), replace=False)]
    cand = np.argwhere(og["pos"] != 0)
    sel = cand[rng.choice(len(cand), size=min(1000 - len(out), len(cand)), replace=False)]
    out += [("pos", int(i), int(j)) for i, j in sel]
    return out


def _pixel_fd(c, k, params):
    """Central differences of pixel k's g . C at steps h and h / 4 for the
    parameters that reach pixel k (on its own oracle object, perturbed in
    place): {param position in the list: (fd_h, fd_h4)}."""
    D = c["objs"][k]
    pos, attr, radii, cm, dm = D.keep[0].reshape(-1, 3), D.keep[1], D.keep[2], D.keep[3], D.keep[4]
    local = {int(g): l for l, g in enumerate(c["subsets"][k])}
    pix = c["px"][k:k + 1]
    gk = c["g"][k]
    L = lambda: float(D.forward(pixels=pix)[0] @ gk)
    out = {}
    for m, (kind, i, j) in enumerate(params):
        if kind in ("pos", "radii", "attr"):
            if i not in local:
                continue
            arr, at = {"pos": (pos, (local[i], j)), "radii": (radii, (local[i],)), "attr": (attr, (local[i],))}[kind]
        else:
            arr, at = (cm, (i, j)) if kind == "colormap" else (dm, (i, 0))
        h = STEPS[kind]
        v0 = arr[at]
        fds = []
        for step in (h, h / 4):
            arr[at] = v0 + step
            lp = L()
            arr[at] = v0 - step
            lm = L()
            arr[at] = v0
            fds.append((lp - lm) / (2 * step))
        out[m] = fds
    return out


def test_c4_finite_differences_on_1000_parameters(c4):
    c = c4
    og = _oracle_adjoint(c)
    params = _params(c, og, np.random.default_rng(11))
    assert len(params) == 1000
    fd_h, fd_q = np.zeros(len(params)), np.zeros(len(params))
    with cf.ThreadPoolExecutor(THREADS) as ex:
        for res in ex.map(lambda k: _pixel_fd(c, k, params), range(NPIX)):
            for m, (a, b) in res.items():
                fd_h[m] += a
                fd_q[m] += b
    gpu = np.array([c["gpu"][kind][i, j] if kind in ("pos", "colormap") else c["gpu"][kind][i]
                    for kind, i, j in params], np.float64)
    kinds = np.array([p[0] for p in params])
    # a sphere boundary or texel kink inside [theta - h, theta + h] makes
    # the two step sizes disagree: those parameters are left out
    smooth = np.abs(fd_h - fd_q) <= 1e-6 + 1e-4 * np.abs(fd_q)
    print(f"\n{(~smooth).sum()} of {len(params)} parameters straddle a discontinuity")
    assert smooth.mean() >= 0.97
    bad_total = 0
    for kind in ("pos", "radii", "attr", "colormap", "densitymap"):
        m = (kinds == kind) & smooth
        a, b = gpu[m], fd_q[m]
        scale = np.abs(b).max()
        rel = np.abs(a - b) / np.maximum(np.abs(b), 1e-2 * scale)
        l2 = float(np.linalg.norm(a - b) / np.linalg.norm(b))
        print(f"{kind}: {m.sum()} parameters, relative L2 {l2:.2e}, max per-parameter {rel.max():.2e}")
        assert l2 < 1e-3, kind
        bad_total += int((rel > 1e-3).sum())
    assert bad_total <= 0.01 * smooth.sum(), f"{bad_total} parameters beyond 1e-3 relative"
