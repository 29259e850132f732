"""HipGrid: FirmCore decomposition of multilayer graphs on one AMD MI355X.

Host side: ``graph`` (generator, reference text format, layer order).
Device engine: ``hipgrid`` (ctypes binding of libhipgrid.so, include/hipgrid.h).
"""
