"""FEM assembly (host, once per mesh): the native direct-to-CSR assembler and
the reference's COO assembler API built on the native element integrals."""
