"""Host-side Python access to the gfx950 Bio-CFD hot-path runtime.

`hip` binds the C ABI (include/biocfd_hip.h), `driver` runs time steps in the
reference driver's call order, `cases` builds the synthetic inputs and
`dumps` reads record files.
"""
