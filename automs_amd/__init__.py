"""AutoMS channel and coverage path on MI355X (gfx950).

Public pieces:
  automs_amd.tracer    build_channel-style tracing through libams.so
  automs_amd.geometry  element grids and placements (go.py helpers)
  automs_amd.scenes    seeded synthetic workloads
"""
__version__ = "0.1.0"
