"""CPU oracle of the adaptive-tree path: TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import
this package, as the parity checker and the timed CPU baseline. The product
package (adaptive_trees_amd) never imports it.
"""
