"""mm3e — host-side Python mirror of the MM3E render-path interface for the MI355X library.

`scene`   : Scene/Object/Prim/Material/Light builder, serialised to the flat IR
            (include/mm3e_scene_ir.h) that libmm3e_hip.so consumes.
`configs` : seeded synthetic workloads.
`hip`     : ctypes bindings over the C ABI (include/mm3e_hip.h).
"""
