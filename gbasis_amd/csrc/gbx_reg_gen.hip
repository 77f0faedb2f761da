// GENERATED by gbasis_amd/codegen.py -- do not edit.
// inputs 8536766e4a5b6837
// Dispatch over the 8 parts of the generated register kernels.
#include "gbx_reg.h"

namespace gbx {

int launch_reg_part0(const gbx_tables* t, int key, const RegArgs& a, int precision, hipStream_t s);
int launch_split_part0(const gbx_tables* t, int key, int split, const RegArgs& a, int precision, hipStream_t s);
int launch_hrr_part0(const gbx_tables* t, int key, const RegArgs& a, hipStream_t s);
int launch_reg_part1(const gbx_tables* t, int key, const RegArgs& a, int precision, hipStream_t s);
int launch_split_part1(const gbx_tables* t, int key, int split, const RegArgs& a, int precision, hipStream_t s);
int launch_hrr_part1(const gbx_tables* t, int key, const RegArgs& a, hipStream_t s);
int launch_reg_part2(const gbx_tables* t, int key, const RegArgs& a, int precision, hipStream_t s);
int launch_split_part2(const gbx_tables* t, int key, int split, const RegArgs& a, int precision, hipStream_t s);
int launch_hrr_part2(const gbx_tables* t, int key, const RegArgs& a, hipStream_t s);
int launch_reg_part3(const gbx_tables* t, int key, const RegArgs& a, int precision, hipStream_t s);
int launch_split_part3(const gbx_tables* t, int key, int split, const RegArgs& a, int precision, hipStream_t s);
int launch_hrr_part3(const gbx_tables* t, int key, const RegArgs& a, hipStream_t s);
int launch_reg_part4(const gbx_tables* t, int key, const RegArgs& a, int precision, hipStream_t s);
int launch_split_part4(const gbx_tables* t, int key, int split, const RegArgs& a, int precision, hipStream_t s);
int launch_hrr_part4(const gbx_tables* t, int key, const RegArgs& a, hipStream_t s);
int launch_reg_part5(const gbx_tables* t, int key, const RegArgs& a, int precision, hipStream_t s);
int launch_split_part5(const gbx_tables* t, int key, int split, const RegArgs& a, int precision, hipStream_t s);
int launch_hrr_part5(const gbx_tables* t, int key, const RegArgs& a, hipStream_t s);
int launch_reg_part6(const gbx_tables* t, int key, const RegArgs& a, int precision, hipStream_t s);
int launch_split_part6(const gbx_tables* t, int key, int split, const RegArgs& a, int precision, hipStream_t s);
int launch_hrr_part6(const gbx_tables* t, int key, const RegArgs& a, hipStream_t s);
int launch_reg_part7(const gbx_tables* t, int key, const RegArgs& a, int precision, hipStream_t s);
int launch_split_part7(const gbx_tables* t, int key, int split, const RegArgs& a, int precision, hipStream_t s);
int launch_hrr_part7(const gbx_tables* t, int key, const RegArgs& a, hipStream_t s);

int launch_reg_generated(const gbx_tables* t, int la, int lb, int lc, int ld, const RegArgs& a, int precision, hipStream_t s) {
  const int key = ((la * 8 + lb) * 8 + lc) * 8 + ld;
  int rc;
  if ((rc = launch_reg_part0(t, key, a, precision, s)) != 1) return rc;
  if ((rc = launch_reg_part1(t, key, a, precision, s)) != 1) return rc;
  if ((rc = launch_reg_part2(t, key, a, precision, s)) != 1) return rc;
  if ((rc = launch_reg_part3(t, key, a, precision, s)) != 1) return rc;
  if ((rc = launch_reg_part4(t, key, a, precision, s)) != 1) return rc;
  if ((rc = launch_reg_part5(t, key, a, precision, s)) != 1) return rc;
  if ((rc = launch_reg_part6(t, key, a, precision, s)) != 1) return rc;
  if ((rc = launch_reg_part7(t, key, a, precision, s)) != 1) return rc;
  return fail(GBX_ERR_UNSUPPORTED, "no fused register kernel for this class");
}

int launch_split_generated(const gbx_tables* t, int la, int lb, int lc, int ld, int k, const RegArgs& a, int precision, hipStream_t s) {
  const int key = ((la * 8 + lb) * 8 + lc) * 8 + ld;
  int rc;
  if ((rc = launch_split_part0(t, key, k, a, precision, s)) != 1) return rc;
  if ((rc = launch_split_part1(t, key, k, a, precision, s)) != 1) return rc;
  if ((rc = launch_split_part2(t, key, k, a, precision, s)) != 1) return rc;
  if ((rc = launch_split_part3(t, key, k, a, precision, s)) != 1) return rc;
  if ((rc = launch_split_part4(t, key, k, a, precision, s)) != 1) return rc;
  if ((rc = launch_split_part5(t, key, k, a, precision, s)) != 1) return rc;
  if ((rc = launch_split_part6(t, key, k, a, precision, s)) != 1) return rc;
  if ((rc = launch_split_part7(t, key, k, a, precision, s)) != 1) return rc;
  return fail(GBX_ERR_UNSUPPORTED, "no split register kernel for this class");
}

int launch_hrr_generated(const gbx_tables* t, int la, int lb, int lc, int ld, const RegArgs& a, hipStream_t s) {
  const int key = ((la * 8 + lb) * 8 + lc) * 8 + ld;
  int rc;
  if ((rc = launch_hrr_part0(t, key, a, s)) != 1) return rc;
  if ((rc = launch_hrr_part1(t, key, a, s)) != 1) return rc;
  if ((rc = launch_hrr_part2(t, key, a, s)) != 1) return rc;
  if ((rc = launch_hrr_part3(t, key, a, s)) != 1) return rc;
  if ((rc = launch_hrr_part4(t, key, a, s)) != 1) return rc;
  if ((rc = launch_hrr_part5(t, key, a, s)) != 1) return rc;
  if ((rc = launch_hrr_part6(t, key, a, s)) != 1) return rc;
  if ((rc = launch_hrr_part7(t, key, a, s)) != 1) return rc;
  return fail(GBX_ERR_UNSUPPORTED, "no register HRR kernel for this class");
}

bool has_hrr_generated(int la, int lb, int lc, int ld) {
  switch (((la * 8 + lb) * 8 + lc) * 8 + ld) {
    case 1105:
    case 1161:
    case 1168:
    case 1169:
    case 649:
    case 650:
    case 657:
    case 1098:
    case 1162:
      return true;
    default:
      return false;
  }
}

bool has_reg_generated(int la, int lb, int lc, int ld) {
  switch (((la * 8 + lb) * 8 + lc) * 8 + ld) {
    case 0:
    case 512:
    case 520:
    case 576:
    case 584:
    case 585:
    case 1024:
    case 1032:
    case 1033:
    case 1040:
    case 1088:
    case 1096:
    case 1097:
    case 1104:
    case 1152:
    case 1160:
    case 640:
    case 648:
    case 656:
    case 1536:
    case 1544:
    case 1545:
    case 1552:
    case 1600:
    case 1608:
    case 1664:
    case 1672:
    case 1728:
    case 704:
    case 712:
    case 1216:
    case 1224:
      return true;
    default:
      return false;
  }
}

int split_count_generated(int la, int lb, int lc, int ld) {
  switch (((la * 8 + lb) * 8 + lc) * 8 + ld) {
    case 1105:
      return 3;
    case 1161:
      return 3;
    case 1168:
      return 2;
    case 1169:
      return 6;
    case 1170:
      return 12;
    case 649:
      return 2;
    case 650:
      return 4;
    case 657:
      return 4;
    case 1098:
      return 4;
    case 1162:
      return 7;
    case 1176:
      return 5;
    case 1553:
      return 2;
    case 1560:
      return 2;
    case 1609:
      return 2;
    case 1616:
      return 2;
    case 1617:
      return 5;
    case 1618:
      return 11;
    case 1624:
      return 4;
    case 1625:
      return 10;
    case 1673:
      return 4;
    case 1680:
      return 3;
    case 1681:
      return 9;
    case 1682:
      return 18;
    case 1688:
      return 8;
    case 1689:
      return 17;
    case 1736:
      return 2;
    case 1737:
      return 6;
    case 1744:
      return 5;
    case 1745:
      return 15;
    case 1746:
      return 32;
    case 1752:
      return 12;
    case 1753:
      return 29;
    case 713:
      return 3;
    case 714:
      return 7;
    case 715:
      return 14;
    case 720:
      return 2;
    case 721:
      return 7;
    case 722:
      return 13;
    case 728:
      return 6;
    case 729:
      return 12;
    case 1225:
      return 5;
    case 1226:
      return 11;
    case 1227:
      return 21;
    case 1232:
      return 4;
    case 1233:
      return 10;
    case 1234:
      return 21;
    case 1240:
      return 8;
    case 1241:
      return 19;
    case 1546:
      return 2;
    case 1610:
      return 5;
    case 1611:
      return 11;
    case 1674:
      return 10;
    case 1675:
      return 21;
    case 1738:
      return 16;
    default:
      return 0;
  }
}

}  // namespace gbx
