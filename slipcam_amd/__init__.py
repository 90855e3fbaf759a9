"""
slipcam_amd: the framecor correlator of slipcam, built for the MI355X (gfx950)

  from slipcam_amd import framecor     # in place of: from slipcam import framecor
"""
