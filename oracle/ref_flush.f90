! Flushes the Fortran runtime's standard output, so that a test can read the
! monitor lines the compiled reference routines print (for example
! srcs/solp_fft_tdma4.f:88-93).  Test infrastructure: linked into
! _ref/libfelref.so only.
subroutine felref_flush() bind(C, name="felref_flush")
  flush(6)
end subroutine felref_flush
