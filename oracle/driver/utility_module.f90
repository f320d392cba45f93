! utility_module.f90 -- the one module of clawpack/classic (commit 1789eeb,
! absent from the reference tree) that the reference's 2-D library uses:
! gauges_module.f90:94 imports get_value_count from it.  Written here from
! its documented purpose (TEST INFRASTRUCTURE ONLY: built with the
! reference's own driver into oracle/_ref).
module utility_module
  implicit none
contains
  ! Number of values on a data-file line: the blank- or comma-separated
  ! items before the first '=' or '#' (the data writers annotate a line of
  ! values with "=: name" or a '#' comment).
  integer function get_value_count(line) result(count)
    character(len=*), intent(in) :: line
    integer :: i, n
    logical :: in_item
    character :: c
    count = 0
    in_item = .false.
    n = len_trim(line)
    do i = 1, n
       c = line(i:i)
       if (c == '=' .or. c == '#') exit
       if (c == ' ' .or. c == ',' .or. c == char(9)) then
          in_item = .false.
       else if (.not. in_item) then
          in_item = .true.
          count = count + 1
       end if
    end do
  end function get_value_count
end module utility_module
