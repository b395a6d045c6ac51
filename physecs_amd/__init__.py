"""physecs_amd: MI355X (gfx950) rigid-body step behind the Physecs Scene::simulate API.

The compute path is libphysecs_gpu.so (HIP kernels + C ABI, include/physecs_gpu.h);
physecs_amd.api binds it, physecs_amd.scenes builds seeded scene descriptors.
"""
__all__ = ["api", "scenes", "build"]
