"""Host-side mirror of PG-Strom's xPU interface for the MI355X library:
ABI structs (abi), kern_data_store builders/parsers (kds), the
kern_expression / session-blob generator (codegen), PostgreSQL hash_any
(hashing), synthetic workloads (datagen) and the ctypes binding of
libxpuhip.so (xpuhip)."""
