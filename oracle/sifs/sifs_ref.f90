! Oracle harness for the SIFS integral file (test infrastructure only).
!
! Built by oracle/sifs/build.sh, linked with colib routines compiled from the reference
! tree: the header writer sifwh (source/colib/colib8.f:3842), the label and dword
! packers plab8 / plab16 (source/colib/colib3.f:1104,1354) and plab10
! (source/colib/colib8.f:4265), and the record writer seqwbf (source/colib/colib2.f).
! The integral records are assembled here the way sife1 / sife2 lay them out
! (source/colib/colib8.f:560-700, :810-990); those two routines copy values through a
! BLAS dcopy that this image lacks, which also makes the reference reader (readmoints)
! unbuildable here.
!
!   sifs_ref SPEC OUT     write OUT from the text description SPEC
!
! SPEC (list-directed): nbft nsym ifmt l1rec n1max l2rec n2max / nbpsy(1:nsym) /
! nenrgy / energy(1:nenrgy) / ietype(1:nenrgy) / n1arrays, then per 1-e array
! "itypea itypeb fcore nval" and nval lines "i j value" / "n2 itypea itypeb" and n2
! lines "i j k l value".  Each 1-e array is split into records of at most n1max values
! (last = 0 inside an array, 1 at its end, 2 after the final array); the 2-e values into
! records of at most n2max (last = 0, then 2 on the final record).
program sifs_ref
  implicit none
  character(len=512) :: spec, outf
  call get_command_argument(1, spec)
  call get_command_argument(2, outf)
  if (len_trim(outf) == 0) stop 'usage: sifs_ref SPEC OUT'
  call write_sifs(trim(spec), trim(outf))
end program sifs_ref

subroutine write_sifs(spec, outf)
  implicit none
  character(len=*), intent(in) :: spec, outf
  integer :: nbft, nsym, ifmt, l1rec, n1max, l2rec, n2max, nenrgy, n1arr, n2
  integer :: info(6), ierr, i, k, a, ta, tb, nval, num, last, p0
  integer, allocatable :: nbpsy(:), ietype(:), lab1(:,:), lab2(:,:), imtype(:), map(:,:)
  double precision, allocatable :: energy(:), v1(:), v2(:), buf(:)
  character(len=80) :: title(1)
  character(len=4), allocatable :: slabel(:)
  character(len=8), allocatable :: bfnlab(:)
  double precision :: fcore
  integer, parameter :: aoints = 21, sp = 22

  open (unit=sp, file=spec, status='old', action='read')
  read (sp, *) nbft, nsym, ifmt, l1rec, n1max, l2rec, n2max
  allocate (nbpsy(nsym), slabel(nsym), bfnlab(nbft))
  read (sp, *) nbpsy
  read (sp, *) nenrgy
  allocate (energy(nenrgy), ietype(nenrgy))
  read (sp, *) energy
  read (sp, *) ietype
  title(1) = 'sifs fixture'
  do i = 1, nsym
    write (slabel(i), '(a1,i3.3)') 's', i
  end do
  do i = 1, nbft
    write (bfnlab(i), '(a2,i6.6)') 'bf', i
  end do
  ! info: fsplit, l1rec, n1max, l2rec, n2max, ifmt
  info = (/ 1, l1rec, n1max, l2rec, n2max, ifmt /)
  allocate (imtype(1), map(nbft, 1))
  open (unit=aoints, file=outf, form='unformatted', status='replace', access='sequential')
  call sifwh(aoints, 1, nsym, nbft, 6, nenrgy, 0, title, nbpsy, slabel, info, bfnlab, &
             ietype, energy, imtype, map, ierr)
  if (ierr /= 0) stop 'sifwh failed'

  allocate (buf(max(l1rec, l2rec)))
  read (sp, *) n1arr
  do a = 1, n1arr
    read (sp, *) ta, tb, fcore, nval
    allocate (v1(max(nval, 1)), lab1(2, max(nval, 1) + 32))
    lab1 = 0
    do k = 1, nval
      read (sp, *) lab1(1, k), lab1(2, k), v1(k)
    end do
    p0 = 0
    do
      num = min(n1max, nval - p0)
      if (p0 + num >= nval) then
        last = 1
        if (a == n1arr) last = 2
      else
        last = 0
      end if
      call enc1(info, num, last, ta, tb, v1(p0 + 1), lab1(1, p0 + 1), fcore, buf)
      call seqwbf(aoints, buf, l1rec)
      p0 = p0 + num
      if (last /= 0) exit
    end do
    deallocate (v1, lab1)
  end do

  read (sp, *) n2, ta, tb
  allocate (v2(max(n2, 1)), lab2(4, max(n2, 1) + 32))
  lab2 = 0
  do k = 1, n2
    read (sp, *) lab2(1, k), lab2(2, k), lab2(3, k), lab2(4, k), v2(k)
  end do
  p0 = 0
  do
    num = min(n2max, n2 - p0)
    last = 0
    if (p0 + num >= n2) last = 2
    call enc2(info, num, last, ta, tb, v2(p0 + 1), lab2(1, p0 + 1), buf)
    call seqwbf(aoints, buf, l2rec)
    p0 = p0 + num
    if (last /= 0) exit
  end do
  close (aoints)
  close (sp)
end subroutine write_sifs

! 1-e record: dword // values(1:num) // fcore // labels packed from word num+3
subroutine enc1(info, num, last, ta, tb, values, labels, fcore, buf)
  implicit none
  integer :: info(6), num, last, ta, tb, labels(2, *), dw(4), k
  double precision :: values(*), fcore, buf(*)
  do k = 1, info(2)
    buf(k) = 0d0
  end do
  dw(1) = num
  dw(2) = num + 3
  dw(3) = (0 * 8 + ta) * 1024 + tb
  dw(4) = info(6) * 4 + last
  call plab16(buf(1), dw, 4)
  do k = 1, num
    buf(1 + k) = values(k)
  end do
  buf(num + 2) = fcore
  if (num == 0) return
  if (info(6) == 0) then
    call plab8(buf(num + 3), labels, 2 * num)
  else
    call plab10(buf(num + 3), labels, 2 * num)
  end if
end subroutine enc1

! 2-e record: dword // values(1:num) // labels packed from word num+2
subroutine enc2(info, num, last, ta, tb, values, labels, buf)
  implicit none
  integer :: info(6), num, last, ta, tb, labels(4, *), dw(4), k
  double precision :: values(*), buf(*)
  do k = 1, info(4)
    buf(k) = 0d0
  end do
  dw(1) = num
  dw(2) = num + 2
  dw(3) = (0 * 8 + ta) * 1024 + tb
  dw(4) = info(6) * 4 + last
  call plab16(buf(1), dw, 4)
  if (num == 0) return
  do k = 1, num
    buf(1 + k) = values(k)
  end do
  if (info(6) == 0) then
    call plab8(buf(num + 2), labels, 4 * num)
  else
    call plab10(buf(num + 2), labels, 4 * num)
  end if
end subroutine enc2
