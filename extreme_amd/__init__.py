"""MI355X-native IE operator apply and FGMRES solve for extrEMe's CIE forward problem."""
