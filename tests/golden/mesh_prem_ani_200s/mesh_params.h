! Proc   1: Header for mesh information to run static solver
! created by the mesher on 12/03/2026, at 03h 50min
 
!:::::::::::::::::::: Input parameters :::::::::::::::::::::::::::
!   Background model     :            prem_ani
!   Dominant period [s]  :  200.0000
!   Elements/wavelength  :    1.5000
!   Courant number       :    0.6000
!   Coarsening levels    :         3
!:::::::::::::::::::::::::::::::::::::::::::::::::::::::::::::::::
 
 integer, parameter ::         npol =         4  !            polynomial order
 integer, parameter ::        nelem =       990  !                   proc. els
 integer, parameter ::       npoint =     24750  !               proc. all pts
 integer, parameter ::    nel_solid =       894  !             proc. solid els
 integer, parameter ::    nel_fluid =        96  !             proc. fluid els
 integer, parameter :: npoint_solid =     22350  !             proc. solid pts
 integer, parameter :: npoint_fluid =      2400  !             proc. fluid pts
 integer, parameter ::  nglob_fluid =      1629  !            proc. flocal pts
 integer, parameter ::     nel_bdry =        36  ! proc. solid-fluid bndry els
 integer, parameter ::        ndisc =        12  !   # disconts in bkgrd model
 integer, parameter ::   nproc_mesh =         1  !        number of processors
 integer, parameter :: lfbkgrdmodel =         8  !   length of bkgrdmodel name
 
!:::::::::::::::::::: Output parameters ::::::::::::::::::::::::::
!   Time step [s]        :    0.1432
!   Min(h/vp),dt/courant :    0.8050    0.9548
!   max(h/vs),T0/wvlngth :  121.8318  133.3333
!   Inner core r_min [km]:  900.3256
!   Max(h) r/ns(icb) [km]:  319.7880
!   Max(h) precalc.  [km]:  356.8605
!:::::::::::::::::::::::::::::::::::::::::::::::::::::::::::::::::
 
