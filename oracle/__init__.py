"""CPU oracle for the DMGS Gaussian-rasterizer hot path — TEST INFRASTRUCTURE ONLY.

Nothing in the product package (``dmgs_amd``) imports this package.  Only the
tests under ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg
of ``bench.py`` use it, and only as the checker / CPU baseline, never as the
thing that is measured or shipped.

Contents (every function cites the reference file:line it restates; the CUDA
rasterizer ``diff-gaussian-rasterization`` is an un-vendored git submodule of the
reference, ``.gitmodules:4-6``, with no pinned commit, so its semantics are
restated from the published upstream algorithm and are marked [U]):

* ``camera``   – view / projection matrices (utils/graphics_utils.py:38-77,
                 scene/cameras.py:48-70) and seeded orbit cameras.
* ``sh``       – real spherical harmonics deg 0-3 (utils/sh_utils.py:26-110).
* ``raster``   – FP64 restatement of the rasterizer forward and the
                 hand-written reverse-mode backward [U]; integer binning
                 (bit-exact, numpy).
* ``raster_autograd`` – an independent differentiable FP64 formulation of the
                 same forward, used to pin ``raster``'s backward by autograd and
                 central finite differences.
* ``binding``  – mesh -> Gaussian binding, stage 3
                 (scene/gaussian_geo_model_finetune.py:396-482) and stage 2
                 (scene/gaussian_geo_model_mlp_flex.py:258-385), FP64 autograd.
* ``loss``     – L1 and SSIM (utils/loss_utils.py:17-62), FP64.

Pinning status (what fixes each function independently of its own formula)
is listed in each module header.  Functions with no independent pin say
"parity unpinned".
"""
