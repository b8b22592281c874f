"""Optional torch.autograd binding of the renderer and its adjoint.

`render(renderer, params, view)` returns the renderer's RGB image (H x W x 3)
as a tensor that is differentiable with respect to the tensors of `params`:

    base_color  (M, 3)   material albedo, renderer materials 0..M-1
    roughness   (M,)     material roughness
    emission    (M, 3)   emitted radiance emissionColor * emissionLuminance
    environment (H, W, 3) environment texels (optional)

The forward pass uploads the parameters through the C API (VKRT_setMaterial,
texture replacement), sets the camera and draws one frame with fresh
accumulation. The backward pass restores the same parameters and view, restarts
the accumulation so the same random streams are replayed, and calls
VKRT_renderBackward (path-replay backpropagation), whose gradients come back as
tensors on the parameters' device. Gradients have no visibility-edge terms.

torch is needed only for this module; the renderer itself does not import it.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from . import Renderer, VKRT_INVALID_INDEX


@dataclass
class SceneParameters:
    base_color: torch.Tensor
    roughness: torch.Tensor
    emission: torch.Tensor
    environment: torch.Tensor | None = None

    @staticmethod
    def from_renderer(r: Renderer, environment: np.ndarray | None = None, device="cpu") -> "SceneParameters":
        """The renderer's current material values (and an environment image) as tensors."""
        m = r.material_count()
        base = np.zeros((m, 3), np.float32)
        rough = np.zeros(m, np.float32)
        emis = np.zeros((m, 3), np.float32)
        for i in range(m):
            mat = r.material(i)
            base[i] = mat.baseColor[:]
            rough[i] = mat.roughness
            emis[i] = np.asarray(mat.emissionColor[:], np.float32) * mat.emissionLuminance
        env = None
        if environment is not None:
            env = torch.tensor(np.asarray(environment, np.float32)[..., :3], device=device)
        return SceneParameters(torch.tensor(base, device=device), torch.tensor(rough, device=device),
                               torch.tensor(emis, device=device), env)

    def tensors(self):
        return [self.base_color, self.roughness, self.emission] + ([self.environment] if self.environment is not None
                                                                 else [])

    def clamp_(self):
        """Projects onto the valid parameter ranges (after an optimiser step)."""
        with torch.no_grad():
            self.base_color.clamp_(0.0, 1.0)
            self.roughness.clamp_(0.0, 1.0)
            self.emission.clamp_(min=0.0)
            if self.environment is not None:
                self.environment.clamp_(min=0.0)


class _Uploaded:
    """What was last written to one renderer, so unchanged values are not re-sent."""

    def __init__(self):
        self.base = self.rough = self.emis = self.environment = None


_uploaded: dict[int, _Uploaded] = {}


def _apply(r: Renderer, base: np.ndarray, rough: np.ndarray, emis: np.ndarray, env: np.ndarray | None):
    st = _uploaded.setdefault(id(r), _Uploaded())
    for i in range(base.shape[0]):
        if (st.base is not None and np.array_equal(st.base[i], base[i]) and st.rough[i] == rough[i]
                and np.array_equal(st.emis[i], emis[i])):
            continue
        mat = r.material(i)
        mat.baseColor[:] = [float(x) for x in base[i]]
        mat.roughness = float(rough[i])
        peak = float(np.max(emis[i]))
        if peak > 0.0:
            mat.emissionColor[:] = [float(x) / peak for x in emis[i]]
            mat.emissionLuminance = peak
        else:
            mat.emissionLuminance = 0.0
        r.set_material(i, mat)
    st.base, st.rough, st.emis = base.copy(), rough.copy(), emis.copy()
    if env is not None and (st.environment is None or not np.array_equal(st.environment, env)):
        r.replace_environment(env)
        st.environment = env.copy()


def _host(t: torch.Tensor | None):
    return None if t is None else t.detach().to("cpu", torch.float32).contiguous().numpy()


class RenderFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, renderer, view, base_color, roughness, emission, environment):
        host = (_host(base_color), _host(roughness), _host(emission), _host(environment))
        _apply(renderer, *host)
        if view is not None:
            renderer.set_camera(*view)
        renderer.start(renderer.width, renderer.height, 0)
        renderer.draw()
        image = renderer.accumulation()[..., :3].copy()
        ctx.renderer, ctx.view, ctx.host = renderer, view, host
        ctx.device = base_color.device
        ctx.has_env = environment is not None
        return torch.from_numpy(image).to(base_color.device)

    @staticmethod
    def backward(ctx, grad_image):
        r = ctx.renderer
        _apply(r, *ctx.host)
        if ctx.view is not None:
            r.set_camera(*ctx.view)
        r.start(r.width, r.height, 0)
        env_shape = ctx.host[3].shape[:2] if ctx.has_env else None
        g = r.render_backward(_host(grad_image), environment_shape=env_shape)
        dev = ctx.device
        genv = torch.from_numpy(g["environment"]).to(dev) if ctx.has_env else None
        return (None, None, torch.from_numpy(g["baseColor"]).to(dev), torch.from_numpy(g["roughness"]).to(dev),
                torch.from_numpy(g["emission"]).to(dev), genv)


def render(renderer: Renderer, params: SceneParameters, view=None) -> torch.Tensor:
    """Differentiable RGB image of the current scene under `params` from `view`
    ((position, target, up, vfov) or None for the current camera)."""
    return RenderFunction.apply(renderer, view, params.base_color, params.roughness, params.emission,
                                params.environment)


def forget(renderer: Renderer):
    """Drops the upload cache of a renderer (call after changing it directly)."""
    _uploaded.pop(id(renderer), None)


__all__ = ["SceneParameters", "RenderFunction", "render", "forget", "VKRT_INVALID_INDEX"]
