"""model3d_amd -- MI355X-native MarchingCubesSearch for model3d solids.

Drop-in for model3d's solid-to-mesh path (model3d/mc.go:45-49): build a
solid with the reference's types, call MarchingCubesSearch, get the mesh.
All compute runs in libmc3d.so (HIP, gfx950); see include/mc3d.h.
"""
from ._lib import (  # noqa: F401
    MC3D_PREC_FP32_GUARD,
    MC3D_PREC_FP64,
    MC3DError,
)
from .mc import (  # noqa: F401
    Contains,
    EncodeSTL,
    MarchingCubes,
    MarchingCubesConj,
    MarchingCubesInterior,
    MarchingCubesSearch,
    MarchingCubesSearchBatch,
    Mesh,
    NewMeshRect,
    Panic,
    WriteSTL,
    case_table,
    interior_map,
)
from .solids import (  # noqa: F401
    CacheSolidBounds,
    Capsule,
    ClipSolid,
    ColliderSolid,
    Cone,
    ConeSlice,
    Cylinder,
    ForceSolidBounds,
    IntersectedSolid,
    JoinedSolid,
    JoinedTransform,
    Matrix3Transform,
    MeshCollider,
    MeshToCollider,
    Mirror,
    NewColliderSolid,
    NewMatrix3Rotation,
    NewRect,
    Rect,
    RotateSolid,
    Rotation,
    Scale,
    ScaleSolid,
    Solid,
    Sphere,
    Subtract,
    SubtractedSolid,
    Torus,
    TransformedSolid,
    TransformSolid,
    Translate,
    TranslateSolid,
    VecScale,
    VecScaleSolid,
    XYZ,
)
