"""CPU oracle for rulinalg's GEMM + factor-and-solve path (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may import this
package, and only as the checker.  The product library never calls it.
"""
