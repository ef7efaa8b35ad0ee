! Test infrastructure, built into _ref/libcdmref.so beside the reference's
! sources: lets the Python checker flush the Fortran runtime's buffered
! standard output (unit 6) while it has stdout redirected, so the reference's
! progress lines end up where the redirection sends them.
subroutine ref_flush_stdout() bind(c, name="ref_flush_stdout")
  flush(6)
end subroutine ref_flush_stdout
