"""gyunity_amd: MI355X-native path-tracing integrator for the Gyunity renderer.

The product is libgyunity.so (HIP kernels for gfx950 behind the C ABI in
include/gyunity.h) and the C++ Integrator in include/gyunity/; this package is
the Python binding used by tests and the benchmark.
"""
from ._lib import (  # noqa: F401
    Batch,
    Context,
    Scene,
    SceneArrays,
    GyError,
    RAY_DTYPE,
    HIT_DTYPE,
    SPHERE_DTYPE,
    MEDIUM_DTYPE,
    TRI_MEDIUM_DTYPE,
    camera,
    camera_matrix,
    generate_scene,
    lib,
    image_read,
    image_to_bytes,
    image_write,
    load_obj,
    orbit_views,
    sppm_params,
    render_params,
    sobol_enum_index,
    sobol_sample,
    halton_sample,
    halton_enum_index,
    GY_SCENE_CORNELL,
    GY_SCENE_ICOSPHERE_GGX,
    GY_SCENE_BATCH_MEMBER,
    GY_SCENE_ADJOINT,
    GY_SCENE_ICOSPHERE_TEXTURED,
    GY_RENDER_MEGAKERNEL,
    GY_RENDER_WAVEFRONT,
    GY_SAMPLER_SOBOL,
    GY_SAMPLER_HALTON,
)
