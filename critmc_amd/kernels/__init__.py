"""Host wrappers of the HIP kernels, one module per model family.

Each mirrors the reference's ``critmc/kernels/<model>.py`` ``run`` interface
with device buffers in place of numpy arrays and a :class:`Chains` record in
place of the numpy ``Generator``.
"""
