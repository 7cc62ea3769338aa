/* shading oracle: filled in below */
