"""curvib_old_amd -- MI355X-native fractional-step core of CURVIB.
"""
