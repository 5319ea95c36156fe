"""CPU oracle for the acoustics3d Helmholtz BEM -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this
package, and only as the checker (or the timed CPU baseline). The product package
acoustic_reliefs_amd never imports it.
"""
