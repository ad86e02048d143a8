"""Learned Newton initializer (drx_mlp_*): inference and training kernels against a float32 PyTorch
reference of the same network, feature extraction against a numpy restatement, and point location with
the learned start against plain location and the FP64-refined oracle.

Tolerances: inference output within 2e-3 absolute (activations are rounded to bf16 between layers; a
rounding flip moves the output by ~1e-4); one training step's gradient within 2% relative norm per
layer of autograd's (the kernel keeps backward deltas and GELU' in bf16), the output-layer bias
gradient (all fp32) within 1e-3; learned-start locations as tight as plain ones (ref coords <= 1e-5 of
the FP64 refinement, located ids identical except where parity_util.explain_mismatches excuses them).
"""
import numpy as np
import pytest
import torch

from devil_ray_amd import synthetic as syn
from devil_ray_amd.synthetic import LOCATION_DTYPE

pytestmark = pytest.mark.gpu

IN, H = 27, 64
# offsets of the fp32 master parameters
W1, B1 = 0, 2048
W2, B2 = 2112, 6208
W3, B3 = 6272, 10368
W4, B4 = 10432, 14528
W5, B5 = 14592, 15616
NPARAM = 15632


def _layers(p):
    return [(p[W1:B1].view(64, 32), p[B1:W2]), (p[W2:B2].view(64, 64), p[B2:W3]), (p[W3:B3].view(64, 64), p[B3:W4]),
            (p[W4:B4].view(64, 64), p[B4:W5]), (p[W5:B5].view(16, 64), p[B5:NPARAM])]


def _bf16(t):
    """bf16 rounding with a straight-through gradient."""
    return t + (t.to(torch.bfloat16).float() - t).detach()


def _gelu(z):
    """x * Phi~(x), Phi~ = cubic smoothstep of the normal CDF over [-2, 2] (the library's GELU)."""
    u = torch.clamp(0.5 * z, -1.0, 1.0)
    return z * (0.5 + u * (0.75 - 0.25 * u * u))


def _ref_forward(p, x):
    """float32 reference: bf16 weights and activations, fp32 accumulate."""
    h = x.float()
    layers = _layers(p)
    for li, (w, b) in enumerate(layers):
        z = h @ _bf16(w).t() + b
        if li < 4:
            h = _bf16(_gelu(z))
        else:
            h = z
    return h[:, :3]


def test_gelu_approximates_normal_cdf():
    z = torch.linspace(-6, 6, 2001, dtype=torch.float64)
    phi = 0.5 * (1 + torch.erf(z / np.sqrt(2.0)))
    assert float((_gelu(z) / torch.where(z == 0, torch.ones_like(z), z) - phi)[z != 0].abs().max()) < 0.029


def _random_features(n, seed):
    g = torch.Generator().manual_seed(seed)
    x = torch.zeros(n, 32)
    x[:, :IN] = torch.rand(n, IN, generator=g) * 2 - 1
    return x.to(torch.bfloat16)


def test_mlp_infer_matches_reference(gpu):
    li = gpu.LearnedInitializer(seed=3)
    p = li.params().clone()
    # make every layer matter: larger last layer than the init's 0.1 scale
    p[W5:B5] *= 10
    p[B1:W2] = 0.1
    li.params().copy_(p)
    packed = li.pack()
    for n in (1, 63, 64, 1000, 4097):
        x = _random_features(n, n)
        base = torch.rand(n, 3)
        out = gpu.LearnedInitializer.infer(packed, x.cuda(), base.cuda()).cpu()
        ref = base + _ref_forward(p.cpu(), x)
        d = (out - ref).abs()
        print(f"  n={n} max {d.max():.3g} mean {d.mean():.3g} |ref| {ref.abs().mean():.3g}", flush=True)
        assert d.max() < 2e-3 and d.mean() < 2e-4


def test_mlp_train_step_gradient(gpu):
    """One Adam step with lr = eps = 1e3 moves every weight by g / (1 + |g|/1e3): the kernel's gradient."""
    n = 256
    li = gpu.LearnedInitializer(seed=5)
    p0 = li.params().clone()
    x = _random_features(n, 11)
    t = torch.rand(n, 3, generator=torch.Generator().manual_seed(12)) * 0.2 - 0.1
    kloss = li.train(x.cuda(), t.cuda(), steps=1, batch=n, lr=1e3, lr_final=-1.0, eps=1e3, beta1=0.9, beta2=0.999)
    torch.cuda.synchronize()
    gk = (p0 - li.params()).cpu()
    gk = gk / (1 - gk.abs() / 1e3)
    pr = p0.cpu().clone().requires_grad_(True)
    y = _ref_forward(pr, x)
    loss = ((y - t) ** 2).sum() / n
    loss.backward()
    gr = pr.grad
    mse = float(loss.detach()) / 3
    print(f"  step mse kernel {float(kloss[0]):.6g} reference {mse:.6g}", flush=True)
    assert abs(float(kloss[0]) - mse) <= 1e-3 * mse
    names = ["W1", "b1", "W2", "b2", "W3", "b3", "W4", "b4", "W5", "b5"]
    offs = [W1, B1, W2, B2, W3, B3, W4, B4, W5, B5, NPARAM]
    for k, name in enumerate(names):
        a, b = gk[offs[k]:offs[k + 1]], gr[offs[k]:offs[k + 1]]
        rel = float((a - b).norm() / b.norm().clamp_min(1e-12))
        print(f"  {name}: |g| {float(b.norm()):.3g} rel err {rel:.3g}", flush=True)
        assert rel < (1e-3 if name == "b5" else 2e-2), f"{name} gradient off by {rel}"
    # padded inputs / outputs stay exactly zero
    assert torch.all(li.params()[W1:B1].view(64, 32)[:, IN:] == 0)
    assert torch.all(li.params()[W5:B5].view(16, 64)[3:] == 0)


def _np_features(mesh, pts, cells):
    """numpy restatement of the feature rows (drx_mlp_features) in float32."""
    p = mesh.order
    n1 = p + 1
    out = np.zeros((len(pts), 32), np.float32)
    base = np.zeros((len(pts), 3), np.float32)
    for r, (pt, e) in enumerate(zip(pts, cells)):
        dofs = mesh.elem_dofs(int(e))
        v = mesh.values[dofs].astype(np.float64)
        org = (0.5 * (v.min(axis=0) + v.max(axis=0))).astype(np.float32)
        rel = (v - org.astype(np.float64)).astype(np.float32)
        corners = []
        for k in range(8):
            i, j, kk = (p if k & 1 else 0), (p if k & 2 else 0), (p if k & 4 else 0)
            corners.append(rel[i + n1 * (j + n1 * kk)])
        c = np.array(corners, np.float32)
        lo, hi = c.min(axis=0), c.max(axis=0)
        ctr = np.float32(0.5) * (lo + hi)
        inv = np.float32(1) / (np.float32(0.5) * (hi - lo))
        q = ((pt.astype(np.float32) - org) - ctr) * inv
        out[r, :3] = q
        out[r, 3:27] = ((c - ctr) * inv).reshape(-1)
        base[r] = np.float32(0.5) * q + np.float32(0.5)
    return out, base


def test_mlp_features(gpu):
    mesh = syn.hex_mesh(4, 3, 0.1, 0.01, 2)
    m = gpu.UnstructuredMesh(mesh)
    pts = syn.query_points(3000, m.bounds(), seed=4)
    loc = m.locate(gpu.to_device(pts.reshape(-1)))
    x, t, b = gpu.LearnedInitializer.features(m, gpu.to_device(pts.reshape(-1)), loc)
    l = loc.cpu().numpy().view(LOCATION_DTYPE)
    ok = l["cell"] >= 0
    assert ok.mean() > 0.99
    fx, fb = _np_features(mesh, pts[ok], l["cell"][ok])
    gx = x.float().cpu().numpy()[ok]
    assert np.all(np.abs(gx - fx) <= np.abs(fx) * 2.0 ** -8 + 1e-6)
    assert np.abs(b.cpu().numpy()[ok] - fb).max() < 1e-5
    assert np.abs(t.cpu().numpy()[ok] - (l["ref"][ok] - fb)).max() < 1e-5
    assert np.all(x.float().cpu().numpy()[~ok] == 0)


def _training_set(gpu, seeds, n_el=6, pts_per=60000):
    xs, ts = [], []
    for s in seeds:
        amp = 0.2 * (s % 5) / 4.0
        mesh = syn.hex_mesh(n_el, 3, amp, 0.01, 100 + s)
        m = gpu.UnstructuredMesh(mesh)
        pts = gpu.to_device(syn.query_points(pts_per, m.bounds(), seed=s).reshape(-1))
        loc = m.locate(pts)
        x, t, _ = gpu.LearnedInitializer.features(m, pts, loc)
        keep = loc.view(torch.int32).view(-1, 4)[:, 0] >= 0
        xs.append(x[keep])
        ts.append(t[keep])
    x, t = torch.cat(xs), torch.cat(ts)
    perm = torch.randperm(x.shape[0], generator=torch.Generator().manual_seed(0)).cuda()
    return x[perm].contiguous(), t[perm].contiguous()


@pytest.fixture(scope="module")
def trained(gpu):
    x, t = _training_set(gpu, range(6))
    li = gpu.LearnedInitializer(seed=1)
    loss = li.train(x, t, steps=600, batch=4096, lr=2e-3, lr_final=2e-5)
    torch.cuda.synchronize()
    return li, loss.cpu().numpy()


def test_mlp_training_reduces_error(trained):
    li, loss = trained
    first, last = loss[:10].mean(), loss[-20:].mean()
    print(f"  mse first {first:.3g} last {last:.3g}", flush=True)
    assert np.all(np.isfinite(loss))
    assert last < first / 5


@pytest.mark.parametrize("which,fallback", [("trained", "affine"), ("trained", "centre"), ("untrained", "affine"),
                                            ("untrained", "centre")])
def test_learned_locate_matches_plain(gpu, orc, trained, which, fallback):
    """Learned start with either plain fallback (the corner-affine start or the SubRef centre) against the
    plain locate: ids as the parity bar allows (parity_util.explain_mismatches), reference coordinates
    within 1e-5 of the FP64 map."""
    from parity_util import explain_mismatches
    li = trained[0] if which == "trained" else gpu.LearnedInitializer(seed=9)
    packed = li.pack()
    mesh = syn.hex_mesh(5, 3, 0.12, 0.01, 777)   # held-out seed
    m = gpu.UnstructuredMesh(mesh)
    pts = syn.query_points(40000, m.bounds(), seed=8)
    dp = gpu.to_device(pts.reshape(-1))
    plain, ps, _ = m.locate(dp, gpu.LocateOptions(start="centre" if fallback == "centre" else "affine"), stats=True)
    learned, ls, _ = m.locate(dp, gpu.LocateOptions(mlp=packed, mlp_fallback=fallback), stats=True)
    torch.cuda.synchronize()
    a, b = plain.cpu().numpy().view(LOCATION_DTYPE), learned.cpu().numpy().view(LOCATION_DTYPE)
    ob = orc.build_mesh_bvh(mesh)
    mism, explained, excess, _ = explain_mismatches(orc, mesh, ob, pts, b["cell"], b["ref"], a["cell"], a["ref"])
    assert explained.all(), f"unexplained mismatches {mism[~explained][:10]} excess {excess[~explained][:10]}"
    assert len(mism) <= max(3, 1e-3 * len(pts))
    both = np.nonzero((a["cell"] == b["cell"]) & (b["cell"] >= 0))[0]
    ex = orc.refine_f64(mesh, pts[both], b["cell"][both], b["ref"][both].astype(np.float64))
    err = np.abs(b["ref"][both].astype(np.float64) - ex).max()
    sp, sl = ps.cpu().numpy(), ls.cpu().numpy()
    print(f"  {which}/{fallback}: located {len(both)} mism {len(mism)} ref err {err:.3g} newton/pt plain "
          f"{sp.mean():.3f} learned {sl.mean():.3f}", flush=True)
    assert err <= 1e-5
    if which == "trained":
        assert sl.mean() < sp.mean(), "a trained initializer should need fewer Newton steps"


def test_learned_locate_launch_forms_identical(gpu, trained):
    """The streamed learned locate gives the same locations, Newton steps and candidate counts for every
    chunk size and launch form (a lane's MLP output does not depend on the other lanes of its round)."""
    packed = trained[0].pack()
    mesh = syn.hex_mesh(5, 3, 0.12, 0.01, 778)
    m = gpu.UnstructuredMesh(mesh)
    for n in (37, 20011):
        dp = gpu.to_device(syn.query_points(n, m.bounds(), seed=9).reshape(-1))
        ref = None
        for launch, chunk in ((1, 1024), (1, 64), (2, 64), (2, 512), (1, 4096)):
            loc, st, ca = m.locate(dp, gpu.LocateOptions(mlp=packed, launch=launch, chunk=chunk), stats=True)
            got = (loc.cpu().numpy(), st.cpu().numpy(), ca.cpu().numpy())
            if ref is None:
                ref = got
            else:
                assert all(np.array_equal(a, b) for a, b in zip(ref, got)), (n, launch, chunk)


def test_learned_integrate_matches_plain(gpu, trained):
    """Volume::integrate with the learned start renders the same image as with the plain start."""
    li = trained[0]
    mesh = syn.hex_mesh(6, 3, 0.1, 0.01, 4242)
    w = h = 40
    m = gpu.UnstructuredMesh(mesh)
    f = gpu.UnstructuredField(mesh)
    rays = gpu.Camera(w, h, 45.0, pos=(0.5, 0.5, 2.5), look_at=(0.5, 0.5, 0.5)).create_rays()
    out = {}
    for name, opts in (("plain", gpu.LocateOptions()), ("learned", gpu.LocateOptions(mlp=li.pack()))):
        vol = gpu.Volume(m, f)
        vol.use_lighting = False
        vol.samples(128)
        vol.table = gpu.ColorTable()
        vol.table.add_point(0.0, (0.5, 0.5, 0.5))
        vol.table.add_point(1.0, (1.0, 0.0, 0.0))
        vol.table.add_alpha(0.0, syn.table_alpha_for_corrected(0.02, 128))
        vol.table.set_smooth(True)
        vol.opts = opts
        parts, n = vol.integrate(rays, tile_width=w, tile_height=h)
        img = gpu.composite(parts, n, w, h)
        torch.cuda.synchronize()
        out[name] = (n, img.cpu().numpy().reshape(-1, 4), vol.last_stats)
    (n0, i0, s0), (n1, i1, s1) = out["plain"], out["learned"]
    d = np.abs(i0 - i1)
    print(f"  partials {n0}/{n1} located {s0.located_samples}/{s1.located_samples} newton {s0.newton_steps}/"
          f"{s1.newton_steps} rgba max diff {d.max():.3g}", flush=True)
    assert n0 == n1 and s0.located_samples == s1.located_samples
    assert d.max() <= 1e-4
