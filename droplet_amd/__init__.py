"""droplet_amd: MI355X-native forward splat renderer and voxelizer."""
from .types import (BackgroundRGBA, Bounds, RenderCamera, SplatData, num_sh_coeffs_per_channel,
                    splat_input_stride)
from .math import mat4_identity, mat4_look_at, vec3
