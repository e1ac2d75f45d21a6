// testcpurectangle -- command-line driver with the arguments and output of the
// reference's testcpurectangle 
// (cudahelmholtz/testcpurectangle.cpp:15-27): cells per axis and the three boundary codes.
#include <cstdlib>

#include "cpu_driver.h"

int main(int argc, char **argv)
{
    if (argc < 7) {
        printf("usage: ./testcpurectangle Nx Ny Nz xbc ybc zbc\n");
        return -1;
    }
    return drv::run_host("testcpurectangle", atoi(argv[1]), atoi(argv[2]), atoi(argv[3]), atoi(argv[4]), atoi(argv[5]),
                        atoi(argv[6]));
}
