"""torch.autograd harness of the path-replay adjoint render (lgpu_render_adjoint).

    r = DiffRenderer(device, scene_desc, camera, 512, 512, spp=64,
                     params=[("texture", 0), ("texture", 1), ("density", 0)])
    img = r.render(albedo, roughness, density, seed=it)   # (H, W, 3) float32 tensor
    loss = ((img - target) ** 2).mean()
    loss.backward()                                        # albedo.grad, roughness.grad, density.grad

Forward uploads the parameter tensors into the committed scene (in-place texture and density
updates) and renders with lgpu_render; backward re-uploads the tensors saved by forward and runs
lgpu_render_adjoint with the incoming image gradient. The adjoint pass uses its own seed (forward
seed + 0x9E3779B9): its samples are independent of the image that produced the loss gradient, so
the gradient of a nonlinear loss carries no correlation bias. Parameter tensors may live on any
torch device: they cross to the host (the renderer's HIP runtime is not torch's), which is a copy
of a few MB per call.
"""
import numpy as np
import torch

from . import abi, gpu

_ADJOINT_SEED = 0x9E3779B9


class DiffRenderer:
    """One differentiable render configuration over a committed scene.

    params: ("texture", texture id) or ("density", medium id) per differentiable tensor, in the
    order render() takes them. Tensor shapes: textures (h, w) or (h, w, channels) as added, grids
    (nz, ny, nx)."""

    def __init__(self, device, scene_desc, camera, width, height, spp, params, max_bounces=6, min_bounces=2):
        self.device, self.camera, self.width, self.height, self.spp = device, camera, width, height, spp
        self.max_bounces, self.min_bounces = max_bounces, min_bounces
        self.scene = gpu.Scene(device)
        self._keep = scene_desc.push(self.scene)
        tex = [o[1] for o in scene_desc.ops if o[0] == "texture"]
        grids = [o[1] for o in scene_desc.ops if o[0] == "medium"]
        self.n_materials = sum(1 for o in scene_desc.ops if o[0] == "material")
        self.tex_sizes = [t.size for t in tex]
        self.tex_shapes = [t.shape for t in tex]
        self.dens_sizes = [np.asarray(g["density"]).size if "density" in g else 0 for g in grids]
        self.dens_shapes = [np.asarray(g["density"]).shape if "density" in g else (0,) for g in grids]
        self.tex_off = np.concatenate([[0], np.cumsum(self.tex_sizes)]).astype(np.int64)
        self.dens_off = np.concatenate([[0], np.cumsum(self.dens_sizes)]).astype(np.int64)
        for kind, i in params:
            if kind == "texture":
                assert 0 <= i < len(tex), "unknown texture id %d" % i
            elif kind == "density":
                assert 0 <= i < len(grids) and self.dens_sizes[i] > 0, "medium %d has no density grid" % i
            else:
                raise ValueError("parameter kind must be 'texture' or 'density'")
        self.params = list(params)
        self.last_stats = None

    def close(self):
        if self.scene is not None:
            self.scene.close()
            self.scene = None

    def _desc(self, film, seed):
        return gpu.render_desc(self.scene, film, self.camera, 0, self.spp, max_bounces=self.max_bounces,
                               min_bounces=self.min_bounces, spectral=abi.SPECTRAL_RGB,
                               estimator=abi.ESTIMATOR_CORRECTED, seed=seed & 0xFFFFFFFF,
                               integrator=abi.INTEGRATOR_VOLPATH)

    def upload(self, tensors):
        for (kind, i), t in zip(self.params, tensors):
            a = t.detach().to("cpu", torch.float32).contiguous().numpy().reshape(-1)
            if kind == "texture":
                assert a.size == self.tex_sizes[i], "texture %d takes %d values" % (i, self.tex_sizes[i])
                self.scene.update_texture(i, a)
            else:
                assert a.size == self.dens_sizes[i], "grid %d takes %d values" % (i, self.dens_sizes[i])
                self.scene.update_density(i, a)

    def forward_image(self, seed):
        film = gpu.Film(self.device, self.width, self.height)
        try:
            self.last_stats = gpu.render(self._desc(film, seed))
            rgb, cnt = film.read()
        finally:
            film.close()
        return (rgb / np.maximum(cnt, 1)[..., None]).astype(np.float32)

    def adjoint(self, d_image, seed):
        """Gradients of sum(d_image * image) per parameter (numpy, the parameters' shapes)."""
        film = gpu.Film(self.device, self.width, self.height)
        try:
            gt, gd, _, st = gpu.render_adjoint(self._desc(film, seed), np.asarray(d_image, dtype=np.float32),
                                               int(self.tex_off[-1]), int(self.dens_off[-1]), self.n_materials)
        finally:
            film.close()
        self.last_stats = st
        out = []
        for kind, i in self.params:
            if kind == "texture":
                out.append(gt[self.tex_off[i]:self.tex_off[i + 1]].reshape(self.tex_shapes[i]))
            else:
                out.append(gd[self.dens_off[i]:self.dens_off[i + 1]].reshape(self.dens_shapes[i]))
        return out

    def render(self, *tensors, seed=0):
        if len(tensors) != len(self.params):
            raise ValueError("render takes %d parameter tensors" % len(self.params))
        return _Render.apply(self, int(seed), *tensors)


class _Render(torch.autograd.Function):
    @staticmethod
    def forward(ctx, r, seed, *tensors):
        r.upload(tensors)
        img = r.forward_image(seed)
        ctx.r, ctx.seed = r, seed
        ctx.save_for_backward(*tensors)
        dev = tensors[0].device if tensors else torch.device("cpu")
        return torch.from_numpy(img).to(dev)

    @staticmethod
    def backward(ctx, grad_img):
        r = ctx.r
        tensors = ctx.saved_tensors
        r.upload(tensors)
        g = grad_img.detach().to("cpu", torch.float32).contiguous().numpy()
        grads = r.adjoint(g, ctx.seed + _ADJOINT_SEED)
        out = [torch.from_numpy(np.ascontiguousarray(x)).to(t.device, t.dtype).reshape(t.shape)
               for x, t in zip(grads, tensors)]
        return (None, None, *out)
