"""CPU oracle for the statevector gate-application path (test infrastructure).

Only tests, the smoke check and the CPU-baseline leg of the benchmark use this
package; the product path never imports it.
"""
