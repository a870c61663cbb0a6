// Planted mistakes for the oracle's pin tests. TEST INFRASTRUCTURE ONLY.
//
// A build with -DNSLO_PLANT=n (n > 0) turns on the one deliberate slip numbered n (a
// dropped term, a wrong sign, a swapped operand) so that a test can show that the
// known-answer, brute-force and closed-form checks of the narrow phase catch it. The
// default build leaves NSLO_PLANT at 0, where every PLANTED(n) is a compile-time false.
#pragma once
#ifndef NSLO_PLANT
#define NSLO_PLANT 0
#endif
#define PLANTED(n) (NSLO_PLANT == (n))
