"""MI355X-native texel-domain screened-Poisson multigrid (gradient-domain texture processing).

The product is libgdtp.so (HIP kernels for gfx950 + native host set-up behind the C ABI in
include/gdtp.h); this package is its Python binding and host-side call surface.
"""
from .api import (METRIC_EMBEDDING, METRIC_UNIFORM, OPERATOR_MASS, OPERATOR_STIFFNESS, OPERATOR_SYSTEM, AddProlongation,
                  Geodesics, HierarchicalSystem, InitializeHierarchy, MultigridSystem, MultiplyByRestriction,
                  LineIntegralConvolution, MultiplyBySystemMatrix_NoReciprocals, Relaxation, RelaxationAndResidual, StiffnessModulation,
                  TextureFilter,
                  TexturedMesh, UpdateLinearSystem, VCycle, embedding_metric, surface_gradient_field, vector_field_metric,
                  Padding, compute_padding, pad_texture_coordinates, pad_image, unpad_image, load_textured_mesh, read_vector,
                  write_vector, read_binary_image, write_binary_image)

__all__ = ["TexturedMesh", "HierarchicalSystem", "MultigridSystem", "InitializeHierarchy", "UpdateLinearSystem", "VCycle",
           "Relaxation", "RelaxationAndResidual", "MultiplyByRestriction", "AddProlongation",
           "MultiplyBySystemMatrix_NoReciprocals", "StiffnessModulation", "OPERATOR_SYSTEM", "OPERATOR_MASS", "OPERATOR_STIFFNESS",
           "METRIC_EMBEDDING", "METRIC_UNIFORM", "TextureFilter", "Geodesics",
           "LineIntegralConvolution", "embedding_metric", "vector_field_metric", "surface_gradient_field",
           "Padding", "compute_padding", "pad_texture_coordinates", "pad_image", "unpad_image", "load_textured_mesh", "read_vector",
           "write_vector", "read_binary_image", "write_binary_image"]
