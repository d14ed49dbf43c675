"""Gravitation kernels.  ``hip1`` is the MI355X (gfx950) HIP kernel; ``_base_`` holds the
plugin contract it implements."""
